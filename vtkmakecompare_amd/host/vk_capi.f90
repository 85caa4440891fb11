! vk_capi.f90 -- ISO_C_BINDING interfaces to the C ABI of include/vtkgpu.h.
!
! Every argument is passed by reference, exactly as the reference's
! Fortran call sites pass it; arrays are assumed-size so any of the
! reference's explicit-shape arrays (halo fields, cellvalue, 1-D metrics)
! can be handed over unchanged.
module vk_capi
  use, intrinsic :: iso_c_binding
  implicit none

  interface
    integer(c_int) function vk_init(device) bind(C, name="vk_init")
      import :: c_int
      integer(c_int), intent(in) :: device
    end function
    integer(c_int) function vk_finalize() bind(C, name="vk_finalize")
      import :: c_int
    end function
    type(c_ptr) function vk_last_error() bind(C, name="vk_last_error")
      import :: c_ptr
    end function
    ! assumed-type: double or float host arrays
    integer(c_int) function vk_register(host, bytes) bind(C, name="vk_register")
      import :: c_int, c_size_t
      type(*), dimension(*) :: host
      integer(c_size_t), intent(in) :: bytes
    end function
    integer(c_int) function vk_upload(host) bind(C, name="vk_upload")
      import :: c_int
      type(*), dimension(*) :: host
    end function
    integer(c_int) function vk_download(host) bind(C, name="vk_download")
      import :: c_int
      type(*), dimension(*) :: host
    end function
    integer(c_int) function vk_synchronize() bind(C, name="vk_synchronize")
      import :: c_int
    end function
    integer(c_int) function vk_extent(host, bytes) bind(C, name="vk_extent")
      import :: c_int, c_size_t
      type(*), dimension(*) :: host
      integer(c_size_t), intent(out) :: bytes
    end function
    integer(c_int) function vk_guard_check(damaged) bind(C, name="vk_guard_check")
      import :: c_int
      integer(c_int), intent(out) :: damaged
    end function

    integer(c_int) function vk_set_bc(u, v, w, p, nx, ny, nz, t) bind(C, name="vk_set_bc")
      import :: c_int, c_double
      real(c_double) :: u(*), v(*), w(*), p(*), t(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_set_convect(u, u_v, v, v_v, w, w_v, dt, nx, ny, nz, &
        dx, dy, dz, dx12, dy12, dz12) bind(C, name="vk_set_convect")
      import :: c_int, c_double
      real(c_double) :: u(*), u_v(*), v(*), v_v(*), w(*), w_v(*)
      real(c_double) :: dt, dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_set_convection_for_scalar(u, v, w, dt, nx, ny, nz, &
        dx, dy, dz, dx12, dy12, dz12, t, u_vt, rhocp) &
        bind(C, name="vk_set_convection_for_scalar")
      import :: c_int, c_double
      real(c_double) :: u(*), v(*), w(*), dt, dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      real(c_double) :: t(*), u_vt(*), rhocp(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_set_conduction(u, u_d, v, v_d, w, w_d, re, dt, nx, ny, nz, &
        dx, dy, dz, dx12, dy12, dz12) bind(C, name="vk_set_conduction")
      import :: c_int, c_double
      real(c_double) :: u(*), u_d(*), v(*), v_d(*), w(*), w_d(*), re, dt
      real(c_double) :: dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_set_diffusion_for_scalar(re, dt, nx, ny, nz, dx, dy, dz, &
        dx12, dy12, dz12, t, u_dt, sc, thc, rhocp) bind(C, name="vk_set_diffusion_for_scalar")
      import :: c_int, c_double
      real(c_double) :: re, dt, dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      real(c_double) :: t(*), u_dt(*), sc, thc(*), rhocp(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_set_immersed_boundary(eta, nx, ny, nz, dt, u_l, v_l, w_l, &
        dx, dy, dz, t_l, cellvalue, u_ta, v_ta, w_ta, l, cons_cell) &
        bind(C, name="vk_set_immersed_boundary")
      import :: c_int, c_double
      real(c_double) :: eta, dt, u_l(*), v_l(*), w_l(*), dx(*), dy(*), dz(*), t_l(*)
      real(c_double) :: cellvalue(*), u_ta(*), v_ta(*), w_ta(*), cons_cell
      integer(c_int) :: nx, ny, nz, l
    end function
    integer(c_int) function vk_set_explicit(u_v, u_pre, u_d, u_a, v_v, v_pre, v_d, v_a, &
        w_v, w_pre, w_d, w_a, u_l, v_l, w_l, nx, ny, nz, l) bind(C, name="vk_set_explicit")
      import :: c_int, c_double
      real(c_double) :: u_v(*), u_pre(*), u_d(*), u_a(*), v_v(*), v_pre(*), v_d(*), v_a(*)
      real(c_double) :: w_v(*), w_pre(*), w_d(*), w_a(*), u_l(*), v_l(*), w_l(*)
      integer(c_int) :: nx, ny, nz, l
    end function
    integer(c_int) function vk_set_explicit_for_tempreture(u_vt, u_dt, u_t, t_pre, t_l, &
        nx, ny, nz, l) bind(C, name="vk_set_explicit_for_tempreture")
      import :: c_int, c_double
      real(c_double) :: u_vt(*), u_dt(*), u_t(*), t_pre(*), t_l(*)
      integer(c_int) :: nx, ny, nz, l
    end function
    integer(c_int) function vk_t_divide(u_t, t, cellvalue, eta, dt, nx, ny, nz) &
        bind(C, name="vk_t_divide")
      import :: c_int, c_double
      real(c_double) :: u_t(*), t(*), cellvalue(*), eta, dt
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_solve_u_temp(u_a, v_a, w_a, nx, ny, nz, dx, dy, dz, &
        dx12, dy12, dz12, re, dt, eta, cellvalue, u, v, w) bind(C, name="vk_solve_u_temp")
      import :: c_int, c_double
      real(c_double) :: u_a(*), v_a(*), w_a(*), dx(*), dy(*), dz(*), dx12(*), dy12(*)
      real(c_double) :: dz12(*), re, dt, eta, cellvalue(*), u(*), v(*), w(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_solve_t(u_t, nx, ny, nz, dx, dy, dz, dx12, dy12, dz12, &
        re, sc, dt, rhocp, thc, cellvalue, eta) bind(C, name="vk_solve_t")
      import :: c_int, c_double
      real(c_double) :: u_t(*), dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      real(c_double) :: re, sc, dt, rhocp(*), thc(*), cellvalue(*), eta
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_mkp(p, u_temp, v_temp, w_temp, dt, dx, dy, dz, dx12, dy12, &
        dz12, re, nx, ny, nz, ompnthreads, u_ta, v_ta, w_ta, u_ep2, v_ep2, w_ep2, divu, &
        u, v, w, xl, zl) bind(C, name="vk_mkp")
      import :: c_int, c_double
      real(c_double) :: p(*), u_temp(*), v_temp(*), w_temp(*), dt, dx(*), dy(*), dz(*)
      real(c_double) :: dx12(*), dy12(*), dz12(*), re, u_ta(*), v_ta(*), w_ta(*)
      real(c_double) :: u_ep2(*), v_ep2(*), w_ep2(*), divu(*), u(*), v(*), w(*), xl, zl
      integer(c_int) :: nx, ny, nz, ompnthreads
    end function
    integer(c_int) function vk_set_levelset(nx, ny, nz, dx, dx12, dy12, dz12, phi0, phi, &
        x, y, z, cellvalue) bind(C, name="vk_set_levelset")
      import :: c_int, c_double
      real(c_double) :: dx(*), dx12(*), dy12(*), dz12(*), phi0(*), phi(*), x(*), y(*), z(*)
      real(c_double) :: cellvalue(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_field_copy(src, dst, nx, ny, nz, interior_only) &
        bind(C, name="vk_field_copy")
      import :: c_int, c_double
      real(c_double) :: src(*), dst(*)
      integer(c_int) :: nx, ny, nz, interior_only
    end function
    integer(c_int) function vk_time_step(u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, &
        u_ta, v_ta, w_ta, u_temp, v_temp, w_temp, divu, cellvalue, rhocp, thc, &
        u_ep2, v_ep2, w_ep2, dx, dy, dz, dx12, dy12, dz12, nx, ny, nz, dt, re, sc, eta, &
        xl, zl, l) bind(C, name="vk_time_step")
      import :: c_int, c_double
      real(c_double) :: u(*), v(*), w(*), p(*), t(*), u_pre(*), v_pre(*), w_pre(*), t_pre(*)
      real(c_double) :: u_ta(*), v_ta(*), w_ta(*), u_temp(*), v_temp(*), w_temp(*), divu(*)
      real(c_double) :: cellvalue(*), rhocp(*), thc(*), u_ep2(*), v_ep2(*), w_ep2(*)
      real(c_double) :: dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      real(c_double) :: dt, re, sc, eta, xl, zl
      integer(c_int) :: nx, ny, nz, l
    end function
    ! the mixed-precision step: t, the history, u_ta, cellvalue, rhoCp and
    ! thc in real(c_float), the rest in real(c_double)
    integer(c_int) function vk_time_step_mixed(u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, &
        u_ta, v_ta, w_ta, u_temp, v_temp, w_temp, divu, cellvalue, rhocp, thc, &
        dx, dy, dz, dx12, dy12, dz12, nx, ny, nz, dt, re, sc, eta, xl, zl, l) &
        bind(C, name="vk_time_step_mixed")
      import :: c_int, c_double, c_float
      real(c_double) :: u(*), v(*), w(*), p(*), u_temp(*), v_temp(*), w_temp(*), divu(*)
      real(c_float) :: t(*), u_pre(*), v_pre(*), w_pre(*), t_pre(*), u_ta(*), v_ta(*), w_ta(*)
      real(c_float) :: cellvalue(*), rhocp(*), thc(*)
      real(c_double) :: dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      real(c_double) :: dt, re, sc, eta, xl, zl
      integer(c_int) :: nx, ny, nz, l
    end function
    ! wall-heat-flux objective of an nsteps run, and its adjoint gradient
    ! with respect to cellvalue(:,:,:,1) (checkpoint every `every` steps)
    integer(c_int) function vk_objective(u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, &
        u_ta, v_ta, w_ta, cellvalue, rhocp, thc, dx, dy, dz, dx12, dy12, dz12, &
        nx, ny, nz, dt, re, sc, eta, xl, zl, nsteps, objective) bind(C, name="vk_objective")
      import :: c_int, c_double
      real(c_double) :: u(*), v(*), w(*), p(*), t(*), u_pre(*), v_pre(*), w_pre(*), t_pre(*)
      real(c_double) :: u_ta(*), v_ta(*), w_ta(*), cellvalue(*), rhocp(*), thc(*)
      real(c_double) :: dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      real(c_double) :: dt, re, sc, eta, xl, zl, objective
      integer(c_int) :: nx, ny, nz, nsteps
    end function
    integer(c_int) function vk_adjoint_gradient(u, v, w, p, t, u_pre, v_pre, w_pre, t_pre, &
        u_ta, v_ta, w_ta, cellvalue, rhocp, thc, dx, dy, dz, dx12, dy12, dz12, &
        nx, ny, nz, dt, re, sc, eta, xl, zl, nsteps, every, objective, grad) &
        bind(C, name="vk_adjoint_gradient")
      import :: c_int, c_double
      real(c_double) :: u(*), v(*), w(*), p(*), t(*), u_pre(*), v_pre(*), w_pre(*), t_pre(*)
      real(c_double) :: u_ta(*), v_ta(*), w_ta(*), cellvalue(*), rhocp(*), thc(*)
      real(c_double) :: dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      real(c_double) :: dt, re, sc, eta, xl, zl, objective, grad(*)
      integer(c_int) :: nx, ny, nz, nsteps, every
    end function
    ! the run-time diagnostics (nse_dominik.f:2519, 2797, 2875); the
    ! results come back in the trailing arguments
    integer(c_int) function vk_check_courant(dt, dx, dy, dz, nx, ny, nz, u, v, w, courant) &
        bind(C, name="vk_check_courant")
      import :: c_int, c_double
      real(c_double) :: dt, dx(*), dy(*), dz(*), u(*), v(*), w(*), courant
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_monitor(nx, ny, nz, l, t, dx, dy, dz, dif, t_all, t_dif, &
        t_before, u, v, w, vel_dif, u_dif, v_dif, w_dif, u_int_before, v_int_before, &
        w_int_before, t_b, dt, int_report, loc) bind(C, name="vk_monitor")
      import :: c_int, c_double
      integer(c_int) :: nx, ny, nz, l, int_report, loc(3)
      real(c_double) :: t(*), dx(*), dy(*), dz(*), dif, t_all, t_dif, t_before(*)
      real(c_double) :: u(*), v(*), w(*), vel_dif, u_dif, v_dif, w_dif
      real(c_double) :: u_int_before, v_int_before, w_int_before, t_b, dt
    end function
    integer(c_int) function vk_check_convergence(y, yl, u, l, dt, re, nx, ny, nz, error_sum) &
        bind(C, name="vk_check_convergence")
      import :: c_int, c_double
      real(c_double) :: y(*), yl, u(*), dt, re, error_sum
      integer(c_int) :: l, nx, ny, nz
    end function
    ! set_levelset.f:202-214: structure.dat into the interior of phi0
    integer(c_int) function vk_readgrid3d(nx, ny, nz, phi0) bind(C, name="vk_readgrid3d")
      import :: c_int, c_double
      integer(c_int) :: nx, ny, nz
      real(c_double) :: phi0(*)
    end function
    ! nse_dominik.f:2726-2793: the field files
    integer(c_int) function vk_outfield(x, y, z, u, v, w, p, nx, ny, nz, tid, t, phi, u_ta, &
        v_ta, w_ta) bind(C, name="vk_outfield")
      import :: c_int, c_double
      real(c_double) :: x(*), y(*), z(*), u(*), v(*), w(*), p(*), tid, t(*), phi(*)
      real(c_double) :: u_ta(*), v_ta(*), w_ta(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_readfield(u, v, w, p, nx, ny, nz, t, phi, u_ta, v_ta, w_ta) &
        bind(C, name="vk_readfield")
      import :: c_int, c_double
      real(c_double) :: u(*), v(*), w(*), p(*), t(*), phi(*), u_ta(*), v_ta(*), w_ta(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_output(u, v, w, p, x, y, z, nx, ny, nz, t, cellvalue, u_ta, &
        v_ta, w_ta, phi, divu) bind(C, name="vk_output")
      import :: c_int, c_double
      real(c_double) :: u(*), v(*), w(*), p(*), x(*), y(*), z(*), t(*), cellvalue(*)
      real(c_double) :: u_ta(*), v_ta(*), w_ta(*), phi(*), divu(*)
      integer(c_int) :: nx, ny, nz
    end function
    ! the pressure projection's components (p2s_fftw2d.f:5,116; mkp.f:223,263,293)
    integer(c_int) function vk_p2s(fp, fpt, nx, ny, nz) bind(C, name="vk_p2s")
      import :: c_int, c_double, c_double_complex
      real(c_double) :: fp(*)
      complex(c_double_complex) :: fpt(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_s2p(fpt, fp, nx, ny, nz) bind(C, name="vk_s2p")
      import :: c_int, c_double, c_double_complex
      complex(c_double_complex) :: fpt(*)
      real(c_double) :: fp(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_spn(nx, ny, nz, ac, kxz, phit) bind(C, name="vk_spn")
      import :: c_int, c_double, c_double_complex
      integer(c_int) :: nx, ny, nz
      real(c_double) :: ac(*), kxz(*)
      complex(c_double_complex) :: phit(*)
    end function
    integer(c_int) function vk_sp0b(nx, ny, nz, ac, phit) bind(C, name="vk_sp0b")
      import :: c_int, c_double, c_double_complex
      integer(c_int) :: nx, ny, nz
      real(c_double) :: ac(*)
      complex(c_double_complex) :: phit(*)
    end function
    integer(c_int) function vk_phibc(nx, ny, nz, phi) bind(C, name="vk_phibc")
      import :: c_int, c_double
      integer(c_int) :: nx, ny, nz
      real(c_double) :: phi(*)
    end function
    ! the batched line-level solvers (tridiag.f:43; periodic_implicit.f:1-93;
    ! nse_dominik.f:1505-1870); line arrays X(nsys, 0:n+1)
    integer(c_int) function vk_ctridiag_batched(a, b, nn, nsys) &
        bind(C, name="vk_ctridiag_batched")
      import :: c_int, c_double, c_double_complex
      real(c_double) :: a(*)
      complex(c_double_complex) :: b(*)
      integer(c_int) :: nn, nsys
    end function
    integer(c_int) function vk_tdma_batched(a, b, c, r, u, n, nsys) &
        bind(C, name="vk_tdma_batched")
      import :: c_int, c_double
      real(c_double) :: a(*), b(*), c(*), r(*), u(*)
      integer(c_int) :: n, nsys
    end function
    integer(c_int) function vk_cyclic_batched(a, b, c, alpha, beta, r, x, n, nsys) &
        bind(C, name="vk_cyclic_batched")
      import :: c_int, c_double
      real(c_double) :: a(*), b(*), c(*), alpha(*), beta(*), r(*), x(*)
      integer(c_int) :: n, nsys
    end function
    integer(c_int) function vk_implicit_solver_batched(sux, nx, dx, dx12, re, dt, ft, fb, eta, &
        cux, nsys) bind(C, name="vk_implicit_solver_batched")
      import :: c_int, c_double
      real(c_double) :: sux(*), dx(*), dx12(*), re, dt, ft(*), fb(*), eta, cux(*)
      integer(c_int) :: nx, nsys
    end function
    integer(c_int) function vk_implicit_uy_batched(s, ny, dy, dy12, re, dt, ft, fb, eta, c, &
        nsys) bind(C, name="vk_implicit_uy_batched")
      import :: c_int, c_double
      real(c_double) :: s(*), dy(*), dy12(*), re, dt, ft(*), fb(*), eta, c(*)
      integer(c_int) :: ny, nsys
    end function
    integer(c_int) function vk_implicit_vy_batched(s, ny, dy, dy12, re, dt, ft, fb, eta, c, &
        nsys) bind(C, name="vk_implicit_vy_batched")
      import :: c_int, c_double
      real(c_double) :: s(*), dy(*), dy12(*), re, dt, ft(*), fb(*), eta, c(*)
      integer(c_int) :: ny, nsys
    end function
    integer(c_int) function vk_implicit_tx_batched(stx, nx, dx, dx12, re, sc, dt, llx, rrx, ft, &
        fb, eta, cux, nsys) bind(C, name="vk_implicit_tx_batched")
      import :: c_int, c_double
      real(c_double) :: stx(*), dx(*), dx12(*), re, sc, dt, llx(*), rrx(*), ft(*), fb(*)
      real(c_double) :: eta, cux(*)
      integer(c_int) :: nx, nsys
    end function
    integer(c_int) function vk_implicit_ty_batched(sty, ny, dy, dy12, re, sc, dt, lly, rry, ft, &
        fb, cuy, eta, nsys) bind(C, name="vk_implicit_ty_batched")
      import :: c_int, c_double
      real(c_double) :: sty(*), dy(*), dy12(*), re, sc, dt, lly(*), rry(*), ft(*), fb(*)
      real(c_double) :: cuy(*), eta
      integer(c_int) :: ny, nsys
    end function
    ! the iterative Poisson solver (nse_dominik.f:1955-2457)
    integer(c_int) function vk_linebyline(nx, ny, nz, u_temp, v_temp, w_temp, dx, dy, dz, &
        dx12, dy12, dz12, dt, p) bind(C, name="vk_linebyline")
      import :: c_int, c_double
      integer(c_int) :: nx, ny, nz
      real(c_double) :: u_temp(*), v_temp(*), w_temp(*), dx(*), dy(*), dz(*), dx12(*)
      real(c_double) :: dy12(*), dz12(*), dt, p(*)
    end function
    integer(c_int) function vk_solve_p(p, u_temp, v_temp, w_temp, dt, dx, dy, dz, dx12, &
        dy12, dz12, nx, ny, nz, res, res_max, res_int, res_var, max_iterations, iterations) &
        bind(C, name="vk_solve_p")
      import :: c_int, c_double, c_ptr
      real(c_double) :: p(*), u_temp(*), v_temp(*), w_temp(*), dt, dx(*), dy(*), dz(*)
      real(c_double) :: dx12(*), dy12(*), dz12(*), res, res_max, res_int, res_var
      integer(c_int) :: nx, ny, nz, iterations
      type(c_ptr), value :: max_iterations
    end function
    ! the adjoint operators (adjoint.f:1,128,185); T_final by address
    integer(c_int) function vk_set_adj_convect(u, u_v, v, v_v, w, w_v, re, dt, nx, ny, nz, &
        dx, dy, dz, dx12, dy12, dz12, u_final, v_final, w_final) &
        bind(C, name="vk_set_adj_convect")
      import :: c_int, c_double
      real(c_double) :: u(*), u_v(*), v(*), v_v(*), w(*), w_v(*), re, dt
      real(c_double) :: dx(*), dy(*), dz(*), dx12(*), dy12(*), dz12(*)
      real(c_double) :: u_final(*), v_final(*), w_final(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_set_adj_convection_for_scalar(u, v, w, re, dt, nx, ny, nz, &
        dx, dy, dz, dx12, dy12, dz12, t, u_vt, rhocp, u_final, v_final, w_final) &
        bind(C, name="vk_set_adj_convection_for_scalar")
      import :: c_int, c_double
      real(c_double) :: u(*), v(*), w(*), re, dt, dx(*), dy(*), dz(*), dx12(*), dy12(*)
      real(c_double) :: dz12(*), t(*), u_vt(*), rhocp(*), u_final(*), v_final(*), w_final(*)
      integer(c_int) :: nx, ny, nz
    end function
    integer(c_int) function vk_set_adj_ns_scalar(dt, nx, ny, nz, dx12, dy12, dz12, t, u_s, &
        v_s, w_s, u_final, v_final, w_final, t_final) bind(C, name="vk_set_adj_ns_scalar")
      import :: c_int, c_double, c_ptr
      real(c_double) :: dt, dx12(*), dy12(*), dz12(*), t(*), u_s(*), v_s(*), w_s(*)
      real(c_double) :: u_final(*), v_final(*), w_final(*)
      type(c_ptr), value :: t_final
      integer(c_int) :: nx, ny, nz
    end function
    ! the legacy VTK files (set_levelset.f:131-198, nse_dominik.f:208-235);
    ! path: a NUL-terminated name or c_null_ptr for the reference's name
    integer(c_int) function vk_vtk_levelset(nx, ny, nz, x, y, z, phi, path) &
        bind(C, name="vk_vtk_levelset")
      import :: c_int, c_double, c_ptr
      integer(c_int) :: nx, ny, nz
      real(c_double) :: x(*), y(*), z(*), phi(*)
      type(c_ptr), value :: path
    end function
    integer(c_int) function vk_vtk_cellvalue(nx, ny, nz, x, y, z, cellvalue, path) &
        bind(C, name="vk_vtk_cellvalue")
      import :: c_int, c_double, c_ptr
      integer(c_int) :: nx, ny, nz
      real(c_double) :: x(*), y(*), z(*), cellvalue(*)
      type(c_ptr), value :: path
    end function
    ! the restart record: the step's whole state and l
    integer(c_int) function vk_write_restart(path, nx, ny, nz, l, u, v, w, p, t, u_pre, v_pre, &
        w_pre, t_pre, u_ta, v_ta, w_ta, rhocp, thc, cellvalue) bind(C, name="vk_write_restart")
      import :: c_int, c_double, c_char
      character(kind=c_char) :: path(*)
      integer(c_int) :: nx, ny, nz, l
      real(c_double) :: u(*), v(*), w(*), p(*), t(*), u_pre(*), v_pre(*), w_pre(*), t_pre(*)
      real(c_double) :: u_ta(*), v_ta(*), w_ta(*), rhocp(*), thc(*), cellvalue(*)
    end function
    integer(c_int) function vk_read_restart(path, nx, ny, nz, l, u, v, w, p, t, u_pre, v_pre, &
        w_pre, t_pre, u_ta, v_ta, w_ta, rhocp, thc, cellvalue) bind(C, name="vk_read_restart")
      import :: c_int, c_double, c_char
      character(kind=c_char) :: path(*)
      integer(c_int) :: nx, ny, nz, l
      real(c_double) :: u(*), v(*), w(*), p(*), t(*), u_pre(*), v_pre(*), w_pre(*), t_pre(*)
      real(c_double) :: u_ta(*), v_ta(*), w_ta(*), rhocp(*), thc(*), cellvalue(*)
    end function
  end interface

  ! whether set_levelset writes levelset.vtk and cellvalue1.vtk as the
  ! reference's does (set_levelset.f:131-198)
  logical, save :: vk_vtk_files = .true.
  character(kind=c_char), target, save :: vk_cellvalue1_name(15) = &
      ['c', 'e', 'l', 'l', 'v', 'a', 'l', 'u', 'e', '1', '.', 'v', 't', 'k', c_null_char]

  ! the field set_adj_NS_scalar reads as T_final, which the reference
  ! leaves an uninitialised local (adjoint.f:190): named by the host with
  ! vk_use_t_final before the first call
  type(c_ptr), save :: vk_t_final = c_null_ptr

contains

  ! Names the (resident) halo field set_adj_NS_scalar reads as T_final.
  subroutine vk_use_t_final(t_final)
    real(c_double), target :: t_final(*)
    vk_t_final = c_loc(t_final)
  end subroutine

  ! Stops with the library's message on a non-zero status (the reference
  ! routines have no status argument, so a failure ends the run).
  subroutine vk_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: n
    if (rc == 0) return
    call c_f_pointer(vk_last_error(), msg, [512])
    text = ' '
    do n = 1, 512
      if (msg(n) == c_null_char) exit
      text(n:n) = msg(n)
    end do
    write (0, '(a,a,i0,a,a)') what, ': vtkgpu status ', rc, ': ', trim(text)
    error stop 2
  end subroutine

  ! Registers a host array of n doubles and copies it to its device mirror.
  subroutine vk_resident(a, n)
    real(c_double) :: a(*)
    integer(c_size_t), intent(in) :: n
    call vk_check(vk_register(a, n * 8_c_size_t), 'vk_register')
    call vk_check(vk_upload(a), 'vk_upload')
  end subroutine

  ! The same for an array of n double complex values.
  subroutine vk_resident_c64(a, n)
    complex(c_double_complex) :: a(*)
    integer(c_size_t), intent(in) :: n
    call vk_check(vk_register(a, n * 16_c_size_t), 'vk_register')
    call vk_check(vk_upload(a), 'vk_upload')
  end subroutine

  ! The same for an array of n floats.
  subroutine vk_resident_f32(a, n)
    real(c_float) :: a(*)
    integer(c_size_t), intent(in) :: n
    call vk_check(vk_register(a, n * 4_c_size_t), 'vk_register')
    call vk_check(vk_upload(a), 'vk_upload')
  end subroutine
end module vk_capi
