! vk_driver.f90 -- a drop-in host for the time loop of program main
! (nse_dominik.f:1-540) on the gfx950 library.
!
! The loop body calls the reference routines by their own names and
! argument lists (vk_reference.f90), in the reference's order
! (nse_dominik.f:290-448); mode 1 calls the fused step instead, mode 2 the
! mixed-precision step (float copies of T, the history, u_ta, cellvalue,
! rhoCp and thc), mode 3 runs nsteps for the wall-heat-flux objective and
! its adjoint gradient with respect to cellvalue(:,:,:,1).  Arrays are
! allocated on the host, made device-resident once, and copied back only
! for the output file.
!
! usage: vk_driver input.bin output.bin
!
! input (stream, native endian):
!   int32   nx, ny, nz, nsteps, mode (0 = routine calls, 1 = fused step,
!           2 = mixed-precision step, 3 = objective and adjoint gradient),
!           levelset (1 = set_levelset from phi0, 0 = cellvalue given),
!           checkpoint interval (mode 3), VTK files (1: set_levelset writes
!           levelset.vtk and cellvalue1.vtk and the driver cellvalue.vtk,
!           as the reference does)
!   float64 dt, Re, Sc, eta, xl, zl, yl, 0
!   float64 dx, dy, dz, dx12, dy12, dz12, x, y, z   (n+2 each)
!   float64 u, v, w, T, phi0                          (halo fields)
!   float64 cellvalue                                 (nx*ny*nz*4)
!   mode 4 only: float64 AC(ny,3), KXZ(0:nx/2,0:nz), A(ny,3), complex B(ny)
! output: u, v, w, p, T, u_ta, v_ta, w_ta, divu, u_temp, v_temp, w_temp
!         (halo fields) and cellvalue; mode 3: the objective, then the
!         gradient (nx*ny*nz); mode 4 (the projection's and the adjoint's
!         component routines by their reference names): P2S of u, the
!         solution phi of SPN + SP0B on P2S of v then S2P and PHIBC,
!         CTRIDIAG's A and B, set_adj_convect's u_v, v_v, w_v,
!         set_adj_convection_for_scalar's u_vT and set_adj_NS_scalar's
!         u_s, v_s, w_s (adjoint fields u_final = T, v_final = w,
!         w_final = v, T_final = phi0)
program vk_driver
  use vk_capi
  implicit none
  integer :: hdr(8), nx, ny, nz, nsteps, mode, lsflag, l, ompnthreads, one, zero
  double precision :: prm(8), dt, re, sc, eta, xl, zl, yl, cons_cell
  double precision :: dif, t_all, t_dif, vel_dif, u_dif, v_dif, w_dif, t_b
  double precision :: u_int_before, v_int_before, w_int_before
  double precision, allocatable, dimension(:) :: dx, dy, dz, dx12, dy12, dz12, x, y, z
  double precision, allocatable, dimension(:,:,:) :: u, v, w, p, t, u_v, v_v, w_v, u_d, v_d, &
      w_d, u_a, v_a, w_a, u_temp, v_temp, w_temp, u_pre, v_pre, w_pre, u_ta, v_ta, w_ta, &
      u_l, v_l, w_l, u_vt, u_dt, u_t, t_pre, t_l, rhocp, thc, divu, u_ep2, v_ep2, w_ep2, &
      phi, phi0
  double precision, allocatable :: cellvalue(:,:,:,:), grad(:,:,:)
  real(c_float), allocatable, dimension(:,:,:) :: tf, u_pref, v_pref, w_pref, t_pref, u_taf, &
      v_taf, w_taf, rhocpf, thcf
  real(c_float), allocatable :: cellvaluef(:,:,:,:)
  double precision :: objective
  double precision, allocatable :: ac(:,:), kxz(:,:), atri(:,:)
  double complex, allocatable :: fpt(:,:,:), fpt2(:,:,:), btri(:)
  integer :: every
  character(len=512) :: fin, fout
  integer(c_size_t) :: nf, nc
  integer :: ios

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (10, file=trim(fin), access='stream', form='unformatted', status='old', iostat=ios)
  if (ios /= 0) error stop 'vk_driver: cannot open the input file'
  read (10) hdr
  read (10) prm
  nx = hdr(1); ny = hdr(2); nz = hdr(3); nsteps = hdr(4); mode = hdr(5); lsflag = hdr(6)
  dt = prm(1); re = prm(2); sc = prm(3); eta = prm(4); xl = prm(5); zl = prm(6)
  yl = prm(7)
  ompnthreads = 1
  one = 1
  zero = 0

  allocate (dx(0:nx+1), dx12(0:nx+1), x(0:nx+1), dy(0:ny+1), dy12(0:ny+1), y(0:ny+1), &
            dz(0:nz+1), dz12(0:nz+1), z(0:nz+1))
  allocate (u(0:nx+1,0:ny+1,0:nz+1))
  allocate (v, w, p, t, u_v, v_v, w_v, u_d, v_d, w_d, u_a, v_a, w_a, u_temp, v_temp, &
            w_temp, u_pre, v_pre, w_pre, u_ta, v_ta, w_ta, u_l, v_l, w_l, u_vt, u_dt, u_t, &
            t_pre, t_l, rhocp, thc, divu, u_ep2, v_ep2, w_ep2, phi, phi0, mold=u)
  allocate (cellvalue(1:nx,1:ny,1:nz,1:4))
  if (size(phi0) /= size(u) .or. lbound(phi0, 1) /= 0) error stop 'vk_driver: allocation'
  read (10) dx, dy, dz, dx12, dy12, dz12, x, y, z
  read (10) u, v, w, t, phi0
  read (10) cellvalue
  if (mode == 4) then
    allocate (ac(1:ny,3), kxz(0:nx/2,0:nz), atri(1:ny,3), btri(1:ny))
    read (10) ac, kxz, atri, btri
  end if
  close (10)

  ! initialisation as program main does it (nse_dominik.f:104-191,246-257)
  p = 0.0d0; u_v = 0.0d0; v_v = 0.0d0; w_v = 0.0d0; u_d = 0.0d0; v_d = 0.0d0; w_d = 0.0d0
  u_a = 0.0d0; v_a = 0.0d0; w_a = 0.0d0; u_temp = 0.0d0; v_temp = 0.0d0; w_temp = 0.0d0
  u_ta = 0.0d0; v_ta = 0.0d0; w_ta = 0.0d0; u_l = 0.0d0; v_l = 0.0d0; w_l = 0.0d0
  u_vt = 0.0d0; u_dt = 0.0d0; u_t = 0.0d0; t_l = 0.0d0; divu = 0.0d0; phi = 0.0d0
  u_ep2 = 0.0d0; v_ep2 = 0.0d0; w_ep2 = 0.0d0
  rhocp = 1.0d0; thc = 1.0d0
  u_pre = u; v_pre = v; w_pre = w; t_pre = t

  call vk_check(vk_init(zero), 'vk_init')
  nf = int(nx + 2, c_size_t) * int(ny + 2, c_size_t) * int(nz + 2, c_size_t)
  nc = int(nx, c_size_t) * int(ny, c_size_t) * int(nz, c_size_t) * 4_c_size_t
  call vk_resident(dx, int(nx + 2, c_size_t)); call vk_resident(dx12, int(nx + 2, c_size_t))
  call vk_resident(x, int(nx + 2, c_size_t))
  call vk_resident(dy, int(ny + 2, c_size_t)); call vk_resident(dy12, int(ny + 2, c_size_t))
  call vk_resident(y, int(ny + 2, c_size_t))
  call vk_resident(dz, int(nz + 2, c_size_t)); call vk_resident(dz12, int(nz + 2, c_size_t))
  call vk_resident(z, int(nz + 2, c_size_t))
  call vk_resident(u, nf); call vk_resident(v, nf); call vk_resident(w, nf)
  call vk_resident(p, nf); call vk_resident(t, nf)
  call vk_resident(u_v, nf); call vk_resident(v_v, nf); call vk_resident(w_v, nf)
  call vk_resident(u_d, nf); call vk_resident(v_d, nf); call vk_resident(w_d, nf)
  call vk_resident(u_a, nf); call vk_resident(v_a, nf); call vk_resident(w_a, nf)
  call vk_resident(u_temp, nf); call vk_resident(v_temp, nf); call vk_resident(w_temp, nf)
  call vk_resident(u_pre, nf); call vk_resident(v_pre, nf); call vk_resident(w_pre, nf)
  call vk_resident(u_ta, nf); call vk_resident(v_ta, nf); call vk_resident(w_ta, nf)
  call vk_resident(u_l, nf); call vk_resident(v_l, nf); call vk_resident(w_l, nf)
  call vk_resident(u_vt, nf); call vk_resident(u_dt, nf); call vk_resident(u_t, nf)
  call vk_resident(t_pre, nf); call vk_resident(t_l, nf)
  call vk_resident(rhocp, nf); call vk_resident(thc, nf); call vk_resident(divu, nf)
  call vk_resident(u_ep2, nf); call vk_resident(v_ep2, nf); call vk_resident(w_ep2, nf)
  call vk_resident(phi, nf); call vk_resident(phi0, nf)
  call vk_resident(cellvalue, nc)

  vk_vtk_files = hdr(8) == 1
  if (lsflag == 1) call set_levelset(nx, ny, nz, dx, dx12, dy12, dz12, phi0, phi, x, y, z, &
                                     cellvalue)
  ! program main's dump of cellvalue(...,1) (nse_dominik.f:208-235)
  if (vk_vtk_files) call vk_check(vk_vtk_cellvalue(nx, ny, nz, x, y, z, cellvalue, c_null_ptr), &
                                  'cellvalue.vtk')

  if (mode == 3) then
    every = max(1, hdr(7))
    allocate (grad(1:nx,1:ny,1:nz))
    grad = 0.0d0
    call vk_resident(grad, int(nx, c_size_t) * int(ny, c_size_t) * int(nz, c_size_t))
    call vk_check(vk_adjoint_gradient(u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, u_ta, v_ta, &
                                      w_ta, cellvalue, rhocp, thc, dx, dy, dz, dx12, dy12, &
                                      dz12, nx, ny, nz, dt, re, sc, eta, xl, zl, nsteps, every, &
                                      objective, grad), 'adjoint gradient')
    call vk_check(vk_download(grad), 'download')
    open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write (11) objective, grad
    close (11)
    call vk_check(vk_finalize(), 'vk_finalize')
    write (6, '(a,i0,a,es23.15)') 'vk_driver: adjoint over ', nsteps, ' steps, J = ', objective
    stop
  end if

  if (mode == 4) then
    allocate (fpt(0:nx/2,0:ny+1,0:nz+1), fpt2(0:nx/2,0:ny+1,0:nz+1))
    fpt = (0.0d0, 0.0d0)
    fpt2 = (0.0d0, 0.0d0)
    nc = int(nx / 2 + 1, c_size_t) * int(ny + 2, c_size_t) * int(nz + 2, c_size_t)
    call vk_resident_c64(fpt, nc); call vk_resident_c64(fpt2, nc)
    call vk_resident(ac, 3_c_size_t * int(ny, c_size_t))
    call vk_resident(kxz, int(nx / 2 + 1, c_size_t) * int(nz + 1, c_size_t))
    call vk_resident(atri, 3_c_size_t * int(ny, c_size_t))
    call vk_resident_c64(btri, int(ny, c_size_t))
    call p2s(u, fpt, nx, ny, nz)
    call p2s(v, fpt2, nx, ny, nz)
    call spn(nx, ny, nz, ac, kxz, fpt2)
    call sp0b(nx, ny, nz, ac, fpt2)
    call s2p(fpt2, phi, nx, ny, nz)
    call phibc(nx, ny, nz, phi)
    call ctridiag(atri, btri, ny)
    call set_adj_convect(u, u_v, v, v_v, w, w_v, re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, &
                         dz12, t, w, v)
    call set_adj_convection_for_scalar(u, v, w, re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, &
                                       dz12, t, u_vt, rhocp, t, w, v)
    call vk_use_t_final(phi0)
    call set_adj_ns_scalar(dt, nx, ny, nz, dx12, dy12, dz12, t, u_l, v_l, w_l, t, w, v)
    call vk_check(vk_download(fpt), 'download'); call vk_check(vk_download(phi), 'download')
    call vk_check(vk_download(atri), 'download'); call vk_check(vk_download(btri), 'download')
    call vk_check(vk_download(u_v), 'download'); call vk_check(vk_download(v_v), 'download')
    call vk_check(vk_download(w_v), 'download'); call vk_check(vk_download(u_vt), 'download')
    call vk_check(vk_download(u_l), 'download'); call vk_check(vk_download(v_l), 'download')
    call vk_check(vk_download(w_l), 'download')
    open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write (11) fpt, phi, atri, btri, u_v, v_v, w_v, u_vt, u_l, v_l, w_l
    close (11)
    call vk_check(vk_finalize(), 'vk_finalize')
    write (6, '(a)') 'vk_driver: component routines done'
    stop
  end if

  if (mode == 2) then
    ! float copies of the fields the mixed step keeps in float
    call vk_check(vk_download(cellvalue), 'download')
    allocate (tf(0:nx+1,0:ny+1,0:nz+1))
    allocate (u_pref, v_pref, w_pref, t_pref, u_taf, v_taf, w_taf, rhocpf, thcf, mold=tf)
    allocate (cellvaluef(1:nx,1:ny,1:nz,1:4))
    tf = real(t, c_float); u_pref = real(u_pre, c_float); v_pref = real(v_pre, c_float)
    w_pref = real(w_pre, c_float); t_pref = real(t_pre, c_float)
    u_taf = real(u_ta, c_float); v_taf = real(v_ta, c_float); w_taf = real(w_ta, c_float)
    rhocpf = real(rhocp, c_float); thcf = real(thc, c_float)
    cellvaluef = real(cellvalue, c_float)
    call vk_resident_f32(tf, nf); call vk_resident_f32(u_pref, nf)
    call vk_resident_f32(v_pref, nf); call vk_resident_f32(w_pref, nf)
    call vk_resident_f32(t_pref, nf); call vk_resident_f32(u_taf, nf)
    call vk_resident_f32(v_taf, nf); call vk_resident_f32(w_taf, nf)
    call vk_resident_f32(rhocpf, nf); call vk_resident_f32(thcf, nf)
    call vk_resident_f32(cellvaluef, nc)
    do l = 1, nsteps
      call vk_check(vk_time_step_mixed(u, v, w, p, tf, u_pref, v_pref, w_pref, t_pref, u_taf, &
                                       v_taf, w_taf, u_temp, v_temp, w_temp, divu, cellvaluef, &
                                       rhocpf, thcf, dx, dy, dz, dx12, dy12, dz12, nx, ny, nz, &
                                       dt, re, sc, eta, xl, zl, l), 'mixed time step')
    end do
    call vk_check(vk_download(tf), 'download'); call vk_check(vk_download(u_taf), 'download')
    call vk_check(vk_download(v_taf), 'download'); call vk_check(vk_download(w_taf), 'download')
    call vk_check(vk_download(cellvaluef), 'download')
    t = real(tf, c_double); u_ta = real(u_taf, c_double); v_ta = real(v_taf, c_double)
    w_ta = real(w_taf, c_double); cellvalue = real(cellvaluef, c_double)
    call vk_check(vk_upload(t), 'upload'); call vk_check(vk_upload(u_ta), 'upload')
    call vk_check(vk_upload(v_ta), 'upload'); call vk_check(vk_upload(w_ta), 'upload')
    call vk_check(vk_upload(cellvalue), 'upload')
  end if

  do l = 1, nsteps
    if (mode == 2) exit
    if (mode == 1) then
      call vk_check(vk_time_step(u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, u_ta, v_ta, w_ta, &
                                 u_temp, v_temp, w_temp, divu, cellvalue, rhocp, thc, u_ep2, &
                                 v_ep2, w_ep2, dx, dy, dz, dx12, dy12, dz12, nx, ny, nz, dt, &
                                 re, sc, eta, xl, zl, l), 'time step')
      cycle
    end if
    ! nse_dominik.f:290-448, report/output branches omitted
    call set_bc(u, v, w, p, nx, ny, nz, t)
    call set_convect(u, u_v, v, v_v, w, w_v, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12)
    call set_convection_for_scalar(u, v, w, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, t, &
                                   u_vt, rhocp)
    call set_conduction(u, u_d, v, v_d, w, w_d, re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, &
                        dz12)
    call set_diffusion_for_scalar(re, dt, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, t, u_dt, &
                                  sc, thc, rhocp)
    call set_immersed_boundary(eta, nx, ny, nz, dt, u_l, v_l, w_l, dx, dy, dz, t_l, &
                               cellvalue, u_ta, v_ta, w_ta, l, cons_cell)
    call set_explicit(u_v, u_pre, u_d, u_a, v_v, v_pre, v_d, v_a, w_v, w_pre, w_d, w_a, u_l, &
                      v_l, w_l, nx, ny, nz, l)
    call set_explicit_for_tempreture(u_vt, u_dt, u_t, t_pre, t_l, nx, ny, nz, l)
    call set_bc(u_a, v_a, w_a, p, nx, ny, nz, t)
    ! u_T = (T + u_T)/(1 + eta cv1 dt), inline in program main (:387-396)
    call vk_check(vk_t_divide(u_t, t, cellvalue, eta, dt, nx, ny, nz), 'T divide')
    call solve_u_temp(u_a, v_a, w_a, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, re, dt, eta, &
                      cellvalue, u, v, w)
    ! u_temp = u_a (:407-417)
    call vk_check(vk_field_copy(u_a, u_temp, nx, ny, nz, zero), 'u_temp copy')
    call vk_check(vk_field_copy(v_a, v_temp, nx, ny, nz, zero), 'v_temp copy')
    call vk_check(vk_field_copy(w_a, w_temp, nx, ny, nz, zero), 'w_temp copy')
    call solve_t(u_t, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, re, sc, dt, rhocp, thc, &
                 cellvalue, eta)
    ! T = u_T on the interior (:429-437)
    call vk_check(vk_field_copy(u_t, t, nx, ny, nz, one), 'T copy')
    call set_bc(u_temp, v_temp, w_temp, p, nx, ny, nz, t)
    call mkp(p, u_temp, v_temp, w_temp, dt, dx, dy, dz, dx12, dy12, dz12, re, nx, ny, nz, &
             ompnthreads, u_ta, v_ta, w_ta, u_ep2, v_ep2, w_ep2, divu, u, v, w, xl, zl)
  end do

  if (mode /= 2) then
    ! the report-interval checks of program main (nse_dominik.f:467-493) on
    ! the final state, reduced on the device: CHECK_COURANT prints, MONITOR
    ! (l = 1, T_pre standing in for T_before) sets its scalars,
    ! CHECK_CONVERGENCE appends to timely_convergence.dat
    call check_courant(dt, dx, dy, dz, nx, ny, nz, u, v, w)
    call monitor(nx, ny, nz, one, t, dx, dy, dz, dif, t_all, t_dif, t_pre, u, v, w, vel_dif, &
                 u_dif, v_dif, w_dif, u_int_before, v_int_before, w_int_before, t_b, dt, zero)
    write (6, '(a,3es25.16)') 'vk_driver: monitor ', t_all, t_dif, u_int_before
    call check_convergence(y, yl, u, nsteps, dt, re, nx, ny, nz)
  end if

  call vk_check(vk_download(u), 'download'); call vk_check(vk_download(v), 'download')
  call vk_check(vk_download(w), 'download'); call vk_check(vk_download(p), 'download')
  call vk_check(vk_download(t), 'download'); call vk_check(vk_download(u_ta), 'download')
  call vk_check(vk_download(v_ta), 'download'); call vk_check(vk_download(w_ta), 'download')
  call vk_check(vk_download(divu), 'download')
  call vk_check(vk_download(u_temp), 'download'); call vk_check(vk_download(v_temp), 'download')
  call vk_check(vk_download(w_temp), 'download')
  call vk_check(vk_download(cellvalue), 'download')
  open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (11) u, v, w, p, t, u_ta, v_ta, w_ta, divu, u_temp, v_temp, w_temp, cellvalue
  close (11)
  call vk_check(vk_finalize(), 'vk_finalize')
  write (6, '(a,i0,a,i0,a,i0,a,i0,a)') 'vk_driver: ', nsteps, ' steps on ', nx, 'x', ny, 'x', &
      nz, ' done'
end program vk_driver
