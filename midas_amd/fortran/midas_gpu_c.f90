! Fortran (iso_c_binding) face of libmidas_gpu.so, declared from
! include/midas_gpu.h.  Used by the drop-in subroutines in midas_dropin.f90;
! it depends on nothing of the reference and compiles on its own.
module midas_gpu_c
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: MIDAS_Q_PI = 1, MIDAS_Q_LITERALS = 2, MIDAS_Q_CEXT_SUM = 4, &
                                 MIDAS_Q_FP32_BLOCKS = 8, MIDAS_Q_ALL = 15
    integer(c_int), parameter :: MIDAS_RX_SD = 0, MIDAS_RX_RF = 1, MIDAS_RX_LB = 2
    integer(c_int), parameter :: MIDAS_INT_SD = 0, MIDAS_INT_LB = 1, MIDAS_INT_UN = 2, MIDAS_INT_GL = 3, &
                                 MIDAS_INT_TR = 4, MIDAS_INT_SM = 5
    integer(c_int), parameter :: MIDAS_SOLVE_FP64 = 0, MIDAS_SOLVE_C32_IR = 1, MIDAS_SOLVE_F16_IR = 2, &
                                 MIDAS_SOLVE_F16X3_IR = 3, MIDAS_SOLVE_BF16_IR = 4, MIDAS_SOLVE_BF16X3_IR = 5

    type, bind(c) :: midas_cells_t
        integer(c_int) :: n
        type(c_ptr) :: x, y, z, sc, eps, kappa, che, znnpp, id
    end type midas_cells_t

    integer(c_int), parameter :: MIDAS_STATUS_CONVERGED = 0, MIDAS_STATUS_NOT_CONVERGED = 1, &
                                 MIDAS_STATUS_FP64_FALLBACK = 2, MIDAS_ERR_NOT_CONVERGED = -4

    type, bind(c) :: midas_solve_opts_t
        integer(c_int) :: max_iter
        real(c_double) :: tol
        integer(c_int) :: nb
        integer(c_int) :: no_fallback
    end type midas_solve_opts_t

    type, bind(c) :: midas_solve_info_t
        integer(c_int) :: info, iterations
        real(c_double) :: berr, anorm, t_factor_ms, t_solve_ms, t_refine_ms
        integer(c_int) :: status, pivot_search
        real(c_double) :: max_l21
    end type midas_solve_info_t

    type, bind(c) :: midas_blocks_t
        integer(c_int) :: nblocks
        type(c_ptr) :: block_start, ext_start, ext_cells
    end type midas_blocks_t

    interface
        integer(c_int) function midas_init(device, ctx) bind(c, name="midas_init")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: ctx
        end function midas_init

        integer(c_int) function midas_finalize(ctx) bind(c, name="midas_finalize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function midas_finalize

        type(c_ptr) function midas_last_error() bind(c, name="midas_last_error")
            import :: c_ptr
        end function midas_last_error

        integer(c_int) function midas_compute_efields_mom(ctx, cells, k0, quirks, ntr, tr_th, tr_ph, nrx, rx_th, &
                                                          rx_ph, mode, opts, s_total, c_ext, c_abs, e, info) &
            bind(c, name="midas_compute_efields_mom")
            import :: c_int, c_ptr, c_double, midas_cells_t, midas_solve_info_t
            type(c_ptr), value :: ctx
            type(midas_cells_t), intent(in) :: cells
            real(c_double), value :: k0
            integer(c_int), value :: quirks, ntr, nrx, mode
            real(c_double), intent(in) :: tr_th(*), tr_ph(*), rx_th(*), rx_ph(*)
            type(c_ptr), value :: opts, s_total, c_ext, c_abs, e
            type(midas_solve_info_t), intent(out) :: info
        end function midas_compute_efields_mom

        integer(c_int) function midas_compute_efields_cbfm(ctx, cells, blocks, k0, quirks, ntr_cbf, th_cbf, ph_cbf, &
                                                           iws, ntr, tr_th, tr_ph, use_aca, kmax, eps_aca, threshold, &
                                                           mode, e, lde, k_out, c_job, ldc, kcap, sr, res_sr, &
                                                           fsr_out, nnz_out, info) &
            bind(c, name="midas_compute_efields_cbfm")
            import :: c_int, c_long, c_ptr, c_double, midas_cells_t, midas_blocks_t, midas_solve_info_t
            type(c_ptr), value :: ctx
            type(midas_cells_t), intent(in) :: cells
            type(midas_blocks_t), intent(in) :: blocks
            real(c_double), value :: k0, eps_aca, threshold, res_sr
            integer(c_int), value :: quirks, ntr_cbf, iws, ntr, use_aca, kmax, mode, kcap, sr
            real(c_double), intent(in) :: th_cbf(*), ph_cbf(*), tr_th(*), tr_ph(*)
            type(c_ptr), value :: e, c_job, fsr_out, nnz_out
            integer(c_long), value :: lde, ldc
            integer(c_int), intent(out) :: k_out(*)
            type(midas_solve_info_t), intent(out) :: info
        end function midas_compute_efields_cbfm

        integer(c_int) function midas_spherical_wave(ctx, cells, k0, quirks, ntr, th, ph, ref32, v, ldv) &
            bind(c, name="midas_spherical_wave")
            import :: c_int, c_long, c_ptr, c_double, midas_cells_t
            type(c_ptr), value :: ctx
            type(midas_cells_t), intent(in) :: cells
            real(c_double), value :: k0
            integer(c_int), value :: quirks, ntr, ref32
            real(c_double), intent(in) :: th(*), ph(*)
            type(c_ptr), value :: v
            integer(c_long), value :: ldv
        end function midas_spherical_wave

        integer(c_int) function midas_near_field(ctx, cells, k0, quirks, e, lde, ntr, nrx, rth, rph, rso, es, lds) &
            bind(c, name="midas_near_field")
            import :: c_int, c_long, c_ptr, c_double, midas_cells_t
            type(c_ptr), value :: ctx, e, es
            type(midas_cells_t), intent(in) :: cells
            real(c_double), value :: k0, rso
            integer(c_int), value :: quirks, ntr, nrx
            integer(c_long), value :: lde, lds
            real(c_double), intent(in) :: rth(*), rph(*)
        end function midas_near_field

        integer(c_int) function midas_scattering_quantities(ctx, s, lds, nrx_tot, nrx, ntr, rx_th, rx_ph, rx_type, &
                                                            rx_w, tr_th, tr_ph, tr_type, tr_w, k0, dm, quirks, &
                                                            c_ext, c_abs, cext_intfields, q, q_av) &
            bind(c, name="midas_scattering_quantities")
            import :: c_int, c_long, c_ptr, c_double
            type(c_ptr), value :: ctx, s, rx_w, tr_w
            integer(c_long), value :: lds
            integer(c_int), value :: nrx_tot, nrx, ntr, rx_type, tr_type, quirks, cext_intfields
            real(c_double), intent(in) :: rx_th(*), rx_ph(*), tr_th(*), tr_ph(*), c_ext(*), c_abs(*)
            real(c_double), value :: k0, dm
            real(c_double), intent(out) :: q(*), q_av(*)
        end function midas_scattering_quantities

        integer(c_int) function midas_scattering_quantities_grid(s, lds, nrx_tot, rx_th, rx_ph, nrx, rx_type, rx_w, &
                                                                 ntr, tr_th, tr_ph, tr_type, k0, dm, quirks, c_ext, &
                                                                 c_abs, cext_intfields, q, q_av) &
            bind(c, name="midas_scattering_quantities_grid")
            import :: c_int, c_long, c_ptr, c_double
            type(c_ptr), value :: s, rx_w
            integer(c_long), value :: lds
            integer(c_int), value :: nrx_tot, nrx, rx_type, ntr, tr_type, quirks, cext_intfields
            real(c_double), intent(in) :: rx_th(*), rx_ph(*), tr_th(*), tr_ph(*), c_ext(*), c_abs(*)
            real(c_double), value :: k0, dm
            real(c_double), intent(out) :: q(*), q_av(*)
        end function midas_scattering_quantities_grid

        integer(c_int) function midas_quad_adaptive(nth, nph, d1, d2, f, ph, jj, a0, b0, res) &
            bind(c, name="midas_quad_adaptive")
            import :: c_int, c_double
            integer(c_int), value :: nth, nph, d1, d2, ph, jj
            real(c_double), intent(in) :: f(*)
            real(c_double), value :: a0, b0
            real(c_double), intent(out) :: res
        end function midas_quad_adaptive

        integer(c_int) function midas_quad_trap2d(nx, ny, f, a, b, c, d, res) bind(c, name="midas_quad_trap2d")
            import :: c_int, c_double
            integer(c_int), value :: nx, ny
            real(c_double), intent(in) :: f(*)
            real(c_double), value :: a, b, c, d
            real(c_double), intent(out) :: res
        end function midas_quad_trap2d

        integer(c_int) function midas_quad_gauss_legendre(n, a, b, x, w) bind(c, name="midas_quad_gauss_legendre")
            import :: c_int, c_double
            integer(c_int), value :: n
            real(c_double), value :: a, b
            real(c_double), intent(out) :: x(*), w(*)
        end function midas_quad_gauss_legendre

        integer(c_int) function midas_extend_blocks(cells, nblocks, block_start, sc, nc_extended, ext_start, &
                                                    ext_cells, ext_cap, adj_start, adj, adj_cap) &
            bind(c, name="midas_extend_blocks")
            import :: c_int, c_long, c_ptr, c_double, midas_cells_t
            type(midas_cells_t), intent(in) :: cells
            integer(c_int), value :: nblocks, nc_extended
            integer(c_int), intent(in) :: block_start(*)
            real(c_double), value :: sc
            integer(c_int), intent(out) :: ext_start(*)
            type(c_ptr), value :: ext_cells, adj_start, adj
            integer(c_long), value :: ext_cap, adj_cap
        end function midas_extend_blocks

        integer(c_int) function midas_partition_sph(cells, navg_cells, best_combin, order, block_size, cap, &
                                                    nblocks, hblock) bind(c, name="midas_partition_sph")
            import :: c_int, c_double, midas_cells_t
            type(midas_cells_t), intent(in) :: cells
            integer(c_int), value :: navg_cells, best_combin, cap
            integer(c_int), intent(out) :: order(*), block_size(*), nblocks
            real(c_double), intent(out) :: hblock
        end function midas_partition_sph

        integer(c_int) function midas_partition_csh(cells, sc, navg_cells, type_s, order, block_size, cap, &
                                                    nblocks, hblock) bind(c, name="midas_partition_csh")
            import :: c_int, c_double, midas_cells_t
            type(midas_cells_t), intent(in) :: cells
            real(c_double), value :: sc
            integer(c_int), value :: navg_cells, type_s, cap
            integer(c_int), intent(out) :: order(*), block_size(*), nblocks
            real(c_double), intent(out) :: hblock
        end function midas_partition_csh

        integer(c_int) function midas_aca_block(ctx, cells, k0, quirks, r0, nr, c0, nc, kmax, eps, u, v, k_out) &
            bind(c, name="midas_aca_block")
            import :: c_int, c_ptr, c_double, midas_cells_t
            type(c_ptr), value :: ctx
            type(midas_cells_t), intent(in) :: cells
            real(c_double), value :: k0, eps
            integer(c_int), value :: quirks, r0, nr, c0, nc, kmax
            type(c_ptr), value :: u, v
            integer(c_int), intent(out) :: k_out
        end function midas_aca_block

        integer(c_int) function midas_aca_matrix(ctx, m, n, a, lda, kmax, eps, u, v, k_out) &
            bind(c, name="midas_aca_matrix")
            import :: c_int, c_long, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: m, n, kmax
            type(c_ptr), value :: a, u, v
            integer(c_long), value :: lda
            real(c_double), value :: eps
            integer(c_int), intent(out) :: k_out
        end function midas_aca_matrix

        integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function c_strlen
    end interface

contains

    ! The library's last error message as a Fortran string.
    function midas_error_message() result(msg)
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: cs(:)
        integer :: i, n
        p = midas_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, cs, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = cs(i)
        end do
    end function midas_error_message

end module midas_gpu_c
