! vk_reference.f90 -- drop-in replacements for the reference routines on
! the time-step path.  Each subroutine has the reference's name, argument
! list and array shapes, so program main's call sites compile unchanged;
! the body forwards to the gfx950 library through vk_capi.  The arrays
! must be device-resident (vk_resident) before the first call.
!
!   SET_BC                      nse_dominik.f:611
!   SET_CONVECT                 nse_dominik.f:757
!   SET_CONVECTION_FOR_SCALAR   nse_dominik.f:834
!   SET_CONDUCTION              nse_dominik.f:888
!   SET_DIFFUSION_FOR_SCALAR    nse_dominik.f:950
!   SET_IMMERSED_BOUNDARY       nse_dominik.f:1011
!   SET_EXPLICIT                nse_dominik.f:1132
!   set_explicit_for_tempreture nse_dominik.f:1291
!   solve_u_temp                nse_dominik.f:1344
!   solve_t                     nse_dominik.f:1465
!   MKP                         mkp.f:9
!   P2S, S2P                    p2s_fftw2d.f:5,116
!   SPN, SP0B, PHIBC            mkp.f:223,263,293
!   CTRIDIAG                    tridiag.f:43
!   set_levelset                set_levelset.f:1
!   set_adj_convect             adjoint.f:1
!   set_adj_convection_for_scalar adjoint.f:128
!   set_adj_NS_scalar           adjoint.f:185
!   linebyline, solve_p         nse_dominik.f:2136,1955

subroutine set_bc(u, v, w, p, nx, ny, nz, t)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: u(0:nx+1,0:ny+1,0:nz+1), v(0:nx+1,0:ny+1,0:nz+1)
  double precision :: w(0:nx+1,0:ny+1,0:nz+1), p(0:nx+1,0:ny+1,0:nz+1)
  double precision :: t(0:nx+1,0:ny+1,0:nz+1)
  call vk_check(vk_set_bc(u, v, w, p, nx, ny, nz, t), 'SET_BC')
end subroutine

subroutine set_convect(u, u_v, v, v_v, w, w_v, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: dt
  double precision :: u(0:nx+1,0:ny+1,0:nz+1), u_v(0:nx+1,0:ny+1,0:nz+1)
  double precision :: v(0:nx+1,0:ny+1,0:nz+1), v_v(0:nx+1,0:ny+1,0:nz+1)
  double precision :: w(0:nx+1,0:ny+1,0:nz+1), w_v(0:nx+1,0:ny+1,0:nz+1)
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_set_convect(u, u_v, v, v_v, w, w_v, dt, nx, ny, nz, dx, dy, dz, &
                               dx12, dy12, dz12), 'SET_CONVECT')
end subroutine

subroutine set_convection_for_scalar(u, v, w, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, &
                                     t, u_vt, rhocp)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: dt
  double precision :: u(0:nx+1,0:ny+1,0:nz+1), v(0:nx+1,0:ny+1,0:nz+1)
  double precision :: w(0:nx+1,0:ny+1,0:nz+1), t(0:nx+1,0:ny+1,0:nz+1)
  double precision :: u_vt(0:nx+1,0:ny+1,0:nz+1), rhocp(0:nx+1,0:ny+1,0:nz+1)
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_set_convection_for_scalar(u, v, w, dt, nx, ny, nz, dx, dy, dz, dx12, &
                                             dy12, dz12, t, u_vt, rhocp), &
                'SET_CONVECTION_FOR_SCALAR')
end subroutine

subroutine set_conduction(u, u_d, v, v_d, w, w_d, re, dt, nx, ny, nz, dx, dy, dz, dx12, &
                          dy12, dz12)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: re, dt
  double precision :: u(0:nx+1,0:ny+1,0:nz+1), u_d(0:nx+1,0:ny+1,0:nz+1)
  double precision :: v(0:nx+1,0:ny+1,0:nz+1), v_d(0:nx+1,0:ny+1,0:nz+1)
  double precision :: w(0:nx+1,0:ny+1,0:nz+1), w_d(0:nx+1,0:ny+1,0:nz+1)
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_set_conduction(u, u_d, v, v_d, w, w_d, re, dt, nx, ny, nz, dx, dy, dz, &
                                  dx12, dy12, dz12), 'SET_CONDUCTION')
end subroutine

subroutine set_diffusion_for_scalar(re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, t, &
                                    u_dt, sc, thc, rhocp)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: re, dt, sc
  double precision :: t(0:nx+1,0:ny+1,0:nz+1), u_dt(0:nx+1,0:ny+1,0:nz+1)
  double precision :: thc(0:nx+1,0:ny+1,0:nz+1), rhocp(0:nx+1,0:ny+1,0:nz+1)
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_set_diffusion_for_scalar(re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, &
                                            dz12, t, u_dt, sc, thc, rhocp), &
                'SET_DIFFUSION_FOR_SCALAR')
end subroutine

subroutine set_immersed_boundary(eta, nx, ny, nz, dt, u_l, v_l, w_l, dx, dy, dz, t_l, &
                                 cellvalue, u_ta, v_ta, w_ta, l, cons_cell)
  use vk_capi
  implicit none
  integer :: nx, ny, nz, l
  double precision :: eta, dt, cons_cell
  double precision :: cellvalue(1:nx,1:ny,1:nz,1:4)
  double precision :: u_l(0:nx+1,0:ny+1,0:nz+1), v_l(0:nx+1,0:ny+1,0:nz+1)
  double precision :: w_l(0:nx+1,0:ny+1,0:nz+1), t_l(0:nx+1,0:ny+1,0:nz+1)
  double precision :: u_ta(0:nx+1,0:ny+1,0:nz+1), v_ta(0:nx+1,0:ny+1,0:nz+1)
  double precision :: w_ta(0:nx+1,0:ny+1,0:nz+1)
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  call vk_check(vk_set_immersed_boundary(eta, nx, ny, nz, dt, u_l, v_l, w_l, dx, dy, dz, &
                                         t_l, cellvalue, u_ta, v_ta, w_ta, l, cons_cell), &
                'SET_IMMERSED_BOUNDARY')
end subroutine

subroutine set_explicit(u_v, u_pre, u_d, u_a, v_v, v_pre, v_d, v_a, w_v, w_pre, w_d, w_a, &
                        u_l, v_l, w_l, nx, ny, nz, l)
  use vk_capi
  implicit none
  integer :: nx, ny, nz, l
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u_v, u_pre, u_d, u_a, v_v, v_pre, &
      v_d, v_a, w_v, w_pre, w_d, w_a, u_l, v_l, w_l
  call vk_check(vk_set_explicit(u_v, u_pre, u_d, u_a, v_v, v_pre, v_d, v_a, w_v, w_pre, w_d, &
                                w_a, u_l, v_l, w_l, nx, ny, nz, l), 'SET_EXPLICIT')
end subroutine

subroutine set_explicit_for_tempreture(u_vt, u_dt, u_t, t_pre, t_l, nx, ny, nz, l)
  use vk_capi
  implicit none
  integer :: nx, ny, nz, l
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u_vt, u_dt, u_t, t_pre, t_l
  call vk_check(vk_set_explicit_for_tempreture(u_vt, u_dt, u_t, t_pre, t_l, nx, ny, nz, l), &
                'set_explicit_for_tempreture')
end subroutine

subroutine solve_u_temp(u_a, v_a, w_a, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, re, dt, &
                        eta, cellvalue, u, v, w)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: re, dt, eta
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u_a, v_a, w_a, u, v, w
  double precision :: cellvalue(1:nx,1:ny,1:nz,1:4)
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_solve_u_temp(u_a, v_a, w_a, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, &
                                re, dt, eta, cellvalue, u, v, w), 'solve_u_temp')
end subroutine

subroutine solve_t(u_t, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, re, sc, dt, rhocp, thc, &
                   cellvalue, eta)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: re, sc, dt, eta
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u_t, rhocp, thc
  double precision :: cellvalue(1:nx,1:ny,1:nz,1:4)
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_solve_t(u_t, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, re, sc, dt, &
                           rhocp, thc, cellvalue, eta), 'solve_t')
end subroutine

subroutine mkp(p, u_temp, v_temp, w_temp, dt, dx, dy, dz, dx12, dy12, dz12, re, nx, ny, nz, &
               ompnthreads, u_ta, v_ta, w_ta, u_ep2, v_ep2, w_ep2, divu, u, v, w, xl, zl)
  use vk_capi
  implicit none
  integer :: nx, ny, nz, ompnthreads
  double precision :: dt, re, xl, zl
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: p, u_temp, v_temp, w_temp, u_ta, &
      v_ta, w_ta, u_ep2, v_ep2, w_ep2, divu, u, v, w
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_mkp(p, u_temp, v_temp, w_temp, dt, dx, dy, dz, dx12, dy12, dz12, re, &
                       nx, ny, nz, ompnthreads, u_ta, v_ta, w_ta, u_ep2, v_ep2, w_ep2, &
                       divu, u, v, w, xl, zl), 'MKP')
end subroutine

subroutine p2s(fp, fpt, nx, ny, nz)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: fp(0:nx+1,0:ny+1,0:nz+1)
  double complex :: fpt(0:nx/2,0:ny+1,0:nz+1)
  call vk_check(vk_p2s(fp, fpt, nx, ny, nz), 'P2S')
end subroutine

subroutine s2p(fpt, fp, nx, ny, nz)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: fp(0:nx+1,0:ny+1,0:nz+1)
  double complex :: fpt(0:nx/2,0:ny+1,0:nz+1)
  call vk_check(vk_s2p(fpt, fp, nx, ny, nz), 'S2P')
end subroutine

subroutine spn(nx, ny, nz, ac, kxz, phit)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: ac(1:ny,3), kxz(0:nx/2,0:nz)
  double complex :: phit(0:nx/2,0:ny+1,0:nz+1)
  call vk_check(vk_spn(nx, ny, nz, ac, kxz, phit), 'SPN')
end subroutine

subroutine sp0b(nx, ny, nz, ac, phit)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: ac(1:ny,3)
  double complex :: phit(0:nx/2,0:ny+1,0:nz+1)
  call vk_check(vk_sp0b(nx, ny, nz, ac, phit), 'SP0B')
end subroutine

subroutine phibc(nx, ny, nz, phi)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: phi(0:nx+1,0:ny+1,0:nz+1)
  call vk_check(vk_phibc(nx, ny, nz, phi), 'PHIBC')
end subroutine

! one system, as the reference calls it: the batched solver with nsys = 1
! has the reference's layout A(NN,3), B(NN)
subroutine ctridiag(a, b, nn)
  use vk_capi
  implicit none
  integer :: nn
  double precision :: a(nn,3)
  double complex :: b(nn)
  call vk_check(vk_ctridiag_batched(a, b, nn, 1), 'CTRIDIAG')
end subroutine

subroutine set_adj_convect(u, u_v, v, v_v, w, w_v, re, dt, nx, ny, nz, dx, dy, dz, dx12, &
                           dy12, dz12, u_final, v_final, w_final)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: re, dt
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u, u_v, v, v_v, w, w_v, u_final, &
      v_final, w_final
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_set_adj_convect(u, u_v, v, v_v, w, w_v, re, dt, nx, ny, nz, dx, dy, dz, &
                                   dx12, dy12, dz12, u_final, v_final, w_final), &
                'set_adj_convect')
end subroutine

subroutine set_adj_convection_for_scalar(u, v, w, re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, &
                                         dz12, t, u_vt, rhocp, u_final, v_final, w_final)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: re, dt
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u, v, w, t, u_vt, rhocp, u_final, &
      v_final, w_final
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_set_adj_convection_for_scalar(u, v, w, re, dt, nx, ny, nz, dx, dy, dz, &
                                                 dx12, dy12, dz12, t, u_vt, rhocp, u_final, &
                                                 v_final, w_final), &
                'set_adj_convection_for_scalar')
end subroutine

! The reference reads T_final from an uninitialised local (adjoint.f:190);
! here it is the field the host named with vk_use_t_final.
subroutine set_adj_ns_scalar(dt, nx, ny, nz, dx12, dy12, dz12, t, u_s, v_s, w_s, u_final, &
                             v_final, w_final)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: dt
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: t, u_s, v_s, w_s, u_final, v_final, &
      w_final
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  if (.not. c_associated(vk_t_final)) then
    write (0, '(a)') 'set_adj_NS_scalar: name T_final with vk_use_t_final first'
    error stop 2
  end if
  call vk_check(vk_set_adj_ns_scalar(dt, nx, ny, nz, dx12, dy12, dz12, t, u_s, v_s, w_s, &
                                     u_final, v_final, w_final, vk_t_final), &
                'set_adj_NS_scalar')
end subroutine

subroutine set_levelset(nx, ny, nz, dx, dx12, dy12, dz12, phi0, phi, x, y, z, cellvalue)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: phi0(0:nx+1,0:ny+1,0:nz+1), phi(0:nx+1,0:ny+1,0:nz+1)
  double precision :: x(0:nx+1), y(0:ny+1), z(0:nz+1)
  double precision :: dx(0:nx+1), dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  double precision :: cellvalue(1:nx,1:ny,1:nz,1:4)
  call vk_check(vk_set_levelset(nx, ny, nz, dx, dx12, dy12, dz12, phi0, phi, x, y, z, &
                                cellvalue), 'set_levelset')
  ! the reference's dumps of phi and cellvalue(...,1) (set_levelset.f:131-198)
  if (vk_vtk_files) then
    call vk_check(vk_vtk_levelset(nx, ny, nz, x, y, z, phi, c_null_ptr), 'levelset.vtk')
    call vk_check(vk_vtk_cellvalue(nx, ny, nz, x, y, z, cellvalue, c_loc(vk_cellvalue1_name)), &
                  'cellvalue1.vtk')
  end if
end subroutine

! CHECK_COURANT (nse_dominik.f:2519-2601): prints the largest Courant
! number as the reference does
subroutine check_courant(dt, dx, dy, dz, nx, ny, nz, u, v, w)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: dt, dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: u(0:nx+1,0:ny+1,0:nz+1), v(0:nx+1,0:ny+1,0:nz+1)
  double precision :: w(0:nx+1,0:ny+1,0:nz+1)
  double precision :: courant
  call vk_check(vk_check_courant(dt, dx, dy, dz, nx, ny, nz, u, v, w, courant), &
                'CHECK_COURANT')
  write(6,*) 'courant max =', courant
end subroutine

! MONITOR (nse_dominik.f:2797-2872), the reference's argument list
subroutine monitor(nx, ny, nz, l, t, dx, dy, dz, dif, t_all, t_dif, t_before, u, v, w, &
                   vel_dif, u_dif, v_dif, w_dif, u_int_before, v_int_before, w_int_before, &
                   t_b, dt, int_report)
  use vk_capi
  implicit none
  integer :: nx, ny, nz, l, int_report
  double precision :: t(0:nx+1,0:ny+1,0:nz+1), t_before(0:nx+1,0:ny+1,0:nz+1)
  double precision :: u(0:nx+1,0:ny+1,0:nz+1), v(0:nx+1,0:ny+1,0:nz+1)
  double precision :: w(0:nx+1,0:ny+1,0:nz+1)
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dif, t_all, t_dif, vel_dif, u_dif, v_dif, w_dif
  double precision :: u_int_before, v_int_before, w_int_before, t_b, dt
  integer :: loc(3)
  call vk_check(vk_monitor(nx, ny, nz, l, t, dx, dy, dz, dif, t_all, t_dif, t_before, u, v, &
                           w, vel_dif, u_dif, v_dif, w_dif, u_int_before, v_int_before, &
                           w_int_before, t_b, dt, int_report, loc), 'MONITOR')
end subroutine

! CHECK_CONVERGENCE (nse_dominik.f:2875-2892): the error sum appended to
! timely_convergence.dat with l*dt, as the reference does
subroutine check_convergence(y, yl, u, l, dt, re, nx, ny, nz)
  use vk_capi
  implicit none
  integer :: l, nx, ny, nz
  double precision :: y(0:ny+1), yl, u(0:nx+1,0:ny+1,0:nz+1), dt, re
  double precision :: error_sum
  call vk_check(vk_check_convergence(y, yl, u, l, dt, re, nx, ny, nz, error_sum), &
                'CHECK_CONVERGENCE')
  open(100, file='timely_convergence.dat', form='formatted', position='APPEND')
  write(100,*) l*dt, error_sum
  close(100)
end subroutine

! READGRID3D (set_levelset.f:202-214): phi0's interior from structure.dat
subroutine readgrid3d(nx, ny, nz, phi0)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: phi0(0:nx+1,0:ny+1,0:nz+1)
  call vk_check(vk_readgrid3d(nx, ny, nz, phi0), 'READGRID3D')
end subroutine

! OUTFIELD (nse_dominik.f:2726-2766): field.u and ./nseoutput2/grid.u
subroutine outfield(x, y, z, u, v, w, p, nx, ny, nz, tid, t, phi, u_ta, v_ta, w_ta)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: x(0:nx+1), y(0:ny+1), z(0:nz+1), tid
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u, v, w, p, t, phi, u_ta, v_ta, w_ta
  call vk_check(vk_outfield(x, y, z, u, v, w, p, nx, ny, nz, tid, t, phi, u_ta, v_ta, w_ta), &
                'OUTFIELD')
end subroutine

! READFIELD (nse_dominik.f:2769-2793): the restart from ./nseoutput2/field.u
subroutine readfield(u, v, w, p, nx, ny, nz, t, phi, u_ta, v_ta, w_ta)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u, v, w, p, t, phi, u_ta, v_ta, w_ta
  call vk_check(vk_readfield(u, v, w, p, nx, ny, nz, t, phi, u_ta, v_ta, w_ta), 'READFIELD')
end subroutine

! OUTPUT (nse_dominik.f:2606-2723): the nseoutput2/ files
subroutine output(u, v, w, p, x, y, z, nx, ny, nz, t, cellvalue, u_ta, v_ta, w_ta, phi, divu)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: x(0:nx+1), y(0:ny+1), z(0:nz+1), cellvalue(1:nx,1:ny,1:nz,1:4)
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u, v, w, p, t, u_ta, v_ta, w_ta, phi, divu
  call vk_check(vk_output(u, v, w, p, x, y, z, nx, ny, nz, t, cellvalue, u_ta, v_ta, w_ta, &
                          phi, divu), 'OUTPUT')
end subroutine

subroutine linebyline(nx, ny, nz, u_temp, v_temp, w_temp, dx, dy, dz, dx12, dy12, dz12, dt, p)
  use vk_capi
  implicit none
  integer :: nx, ny, nz
  double precision :: dt
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: u_temp, v_temp, w_temp, p
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_linebyline(nx, ny, nz, u_temp, v_temp, w_temp, dx, dy, dz, dx12, dy12, &
                              dz12, dt, p), 'linebyline')
end subroutine

! solve_p (nse_dominik.f:1955-2021): the iteration on the device, the
! reference's messages on unit 6
subroutine solve_p(p, u_temp, v_temp, w_temp, dt, dx, dy, dz, dx12, dy12, dz12, nx, ny, nz, &
                   res, res_max, res_int, res_var)
  use vk_capi
  implicit none
  integer :: nx, ny, nz, iterations
  double precision :: dt, res, res_max, res_int, res_var
  double precision, dimension(0:nx+1,0:ny+1,0:nz+1) :: p, u_temp, v_temp, w_temp
  double precision :: dx(0:nx+1), dy(0:ny+1), dz(0:nz+1)
  double precision :: dx12(0:nx+1), dy12(0:ny+1), dz12(0:nz+1)
  call vk_check(vk_solve_p(p, u_temp, v_temp, w_temp, dt, dx, dy, dz, dx12, dy12, dz12, nx, &
                           ny, nz, res, res_max, res_int, res_var, c_null_ptr, iterations), &
                'solve_p')
  write (*, *) 'convergence of poisson =', res_max
  write (*, *) 'res = ', res
  write (*, *) 'res_maximum = ', res_max
  write (*, *) 'res_integral = ', res_int
  write (*, *) 'res_var = ', res_var
end subroutine
