! C-callable driver around the reference's own hot-path routines.
!
! TEST INFRASTRUCTURE ONLY.  It is compiled by oracle/Makefile.ref together
! with the unmodified reference sources (Initialization.f90, Common_variables.f90,
! SetCellsParams.f90, Green_functions.f90, Incident_Field.f90,
! Calcul_ACA_matrices.f90, getParameters_CBFM.f90, read in place under the
! reference tree) into
! oracle/_ref/libmidas_ref.so, which tests use to pin the oracle.  Nothing in
! the product loads it.
!
! Each wrapper takes plain arrays (SoA cells, interleaved complex128), builds
! the reference's `Cell` / `Dipole` records, sets the module globals the
! routine reads, and calls the routine unchanged.
module midas_ref_driver
    use iso_c_binding
    use Initialization
    use common_variables
    implicit none
contains

    subroutine build_cells(n, id, x, y, z, sc, kappa, che, znnpp, cells)
        integer, intent(in) :: n
        integer(c_int), intent(in) :: id(n)
        real(c_double), intent(in) :: x(n), y(n), z(n), sc(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n), znnpp(n)
        type(Cell), intent(out) :: cells(n)
        integer :: c
        do c = 1, n
            cells(c)%n_cell = id(c)
            cells(c)%n_block = 1
            cells(c)%n_diel = 1
            cells(c)%Dlamb_n = 0
            cells(c)%lambda_n = 0d0
            cells(c)%Sc = sc(c)
            cells(c)%Xc = x(c)
            cells(c)%Yc = y(c)
            cells(c)%Zc = z(c)
            cells(c)%a_n = 0d0
            cells(c)%kappa_n = kappa(c)
            cells(c)%m_n = (0d0, 0d0)
            cells(c)%Eps_n = che(c) + 1d0
            cells(c)%Che_n = che(c)
            cells(c)%Znnpp = znnpp(c)
        end do
    end subroutine build_cells

    subroutine ref_set_globals(k0, nbc, ntr, homogs_in, nb_it, eps_aca) bind(c, name="ref_set_globals")
        real(c_double), value :: k0, eps_aca
        integer(c_int), value :: nbc, ntr, homogs_in, nb_it
        k_0 = k0
        Nbc = nbc
        NTr = ntr
        homogs = homogs_in
        Nb_it_max = nb_it
        Epsilon_ACA = eps_aca
        vrb_ACA = 0
        rank = 0
        nber_procs = 1
    end subroutine ref_set_globals

    ! SetCellsParams (SetCellsParams.f90:1-50)
    subroutine ref_cell_params(n, sc, eps, a_out, che_out, znnpp_out, kappa_out) bind(c, name="ref_cell_params")
        integer(c_int), value :: n
        real(c_double), intent(in) :: sc(n)
        complex(c_double_complex), intent(in) :: eps(n)
        real(c_double), intent(out) :: a_out(n), kappa_out(n)
        complex(c_double_complex), intent(out) :: che_out(n), znnpp_out(n)
        interface
            subroutine SetCellsParams(SimScatterer, Cells, UCells)
                use Initialization
                use common_variables
                type(Scatterer), intent(inout) :: SimScatterer
                type(Cell), dimension(Nbc), intent(inout) :: Cells
                type(Cell), dimension(:), allocatable, intent(out) :: UCells
            end subroutine SetCellsParams
        end interface
        type(Scatterer) :: sim
        type(Cell), allocatable :: cells(:), ucells(:)
        integer :: c
        Nbc = n
        allocate(cells(n))
        do c = 1, n
            cells(c)%n_cell = c
            cells(c)%Sc = sc(c)
            cells(c)%Eps_n = eps(c)
            cells(c)%Xc = 0d0; cells(c)%Yc = 0d0; cells(c)%Zc = 0d0
        end do
        call SetCellsParams(sim, cells, ucells)
        do c = 1, n
            a_out(c) = ucells(c)%a_n
            che_out(c) = ucells(c)%Che_n
            znnpp_out(c) = ucells(c)%Znnpp
            kappa_out(c) = ucells(c)%Kappa_n
        end do
    end subroutine ref_cell_params

    ! Green_s_tr_partial (Green_functions.f90:96-173)
    subroutine ref_block_fill(n1, id1, x1, y1, z1, n2, id2, x2, y2, z2, kappa2, che2, znnpp2, Z) &
            bind(c, name="ref_block_fill")
        integer(c_int), value :: n1, n2
        integer(c_int), intent(in) :: id1(n1), id2(n2)
        real(c_double), intent(in) :: x1(n1), y1(n1), z1(n1), x2(n2), y2(n2), z2(n2), kappa2(n2)
        complex(c_double_complex), intent(in) :: che2(n2), znnpp2(n2)
        complex(c_double_complex), intent(out) :: Z(3*n1, 3*n2)
        type(Cell), allocatable :: c1(:), c2(:)
        real(c_double), allocatable :: zero1(:)
        complex(c_double_complex), allocatable :: zc1(:)
        allocate(c1(n1), c2(n2), zero1(n1), zc1(n1))
        zero1 = 0d0; zc1 = (0d0, 0d0)
        call build_cells(n1, id1, x1, y1, z1, zero1, zero1, zc1, zc1, c1)
        call build_cells(n2, id2, x2, y2, z2, kappa2, kappa2, che2, znnpp2, c2)
        call Green_s_tr_partial(n1, c1, n2, c2, Z)
    end subroutine ref_block_fill

    ! Green_s_tr_total (Green_functions.f90:21-94)
    subroutine ref_total_fill(n, x, y, z, kappa, che, znnpp, zmat) bind(c, name="ref_total_fill")
        integer(c_int), value :: n
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n), znnpp(n)
        complex(c_double_complex), intent(out) :: zmat(3*n, 3*n)
        type(Cell), allocatable :: cc(:)
        integer(c_int), allocatable :: id(:)
        integer :: c
        allocate(cc(n), id(n))
        id = [(c, c = 1, n)]
        call build_cells(n, id, x, y, z, kappa, kappa, che, znnpp, cc)
        Nbc = n
        call Green_s_tr_total(cc, zmat)
    end subroutine ref_total_fill

    ! computeBlockRow / computeBlockCol (Green_functions.f90:905-1012, 1119-1225);
    ! grow/gcol and the cursors are the reference's 1-based global indexes.
    subroutine ref_block_row(n, x, y, z, kappa, che, grow, col_inf, ncols, row) bind(c, name="ref_block_row")
        integer(c_int), value :: n, grow, col_inf, ncols
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n)
        complex(c_double_complex), intent(out) :: row(1, ncols)
        type(Cell), allocatable :: cc(:)
        integer(c_int), allocatable :: id(:)
        integer :: c
        allocate(cc(n), id(n))
        id = [(c, c = 1, n)]
        call build_cells(n, id, x, y, z, kappa, kappa, che, che, cc)
        Nbc = n
        call computeBlockRow(cc, grow, col_inf, ncols, row)
    end subroutine ref_block_row

    subroutine ref_block_col(n, x, y, z, kappa, che, row_inf, nrows, gcol, col) bind(c, name="ref_block_col")
        integer(c_int), value :: n, row_inf, nrows, gcol
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n)
        complex(c_double_complex), intent(out) :: col(nrows, 1)
        type(Cell), allocatable :: cc(:)
        integer(c_int), allocatable :: id(:)
        integer :: c
        allocate(cc(n), id(n))
        id = [(c, c = 1, n)]
        call build_cells(n, id, x, y, z, kappa, kappa, che, che, cc)
        Nbc = n
        call computeBlockCol(cc, row_inf, nrows, gcol, col)
    end subroutine ref_block_col

    ! Incident_Field (Incident_Field.f90:1-63); transmitters t0..t1 are 1-based.
    subroutine ref_incident_field(n, x, y, z, ntr, th, ph, t0, t1, E) bind(c, name="ref_incident_field")
        integer(c_int), value :: n, ntr, t0, t1
        real(c_double), intent(in) :: x(n), y(n), z(n), th(ntr), ph(ntr)
        complex(c_double_complex), intent(out) :: E(3*n, 2*(t1-t0+1))
        type(Cell), allocatable :: cc(:)
        type(Dipole), allocatable :: tr(:)
        integer(c_int), allocatable :: id(:)
        real(c_double), allocatable :: zr(:)
        complex(c_double_complex), allocatable :: zc(:)
        integer :: c
        allocate(cc(n), id(n), tr(ntr), zr(n), zc(n))
        id = [(c, c = 1, n)]; zr = 0d0; zc = (0d0, 0d0)
        call build_cells(n, id, x, y, z, zr, zr, zc, zc, cc)
        do c = 1, ntr
            tr(c)%theta = th(c)
            tr(c)%phi = ph(c)
        end do
        call Incident_Field(1, n, cc, ntr, tr, t0, t1, E)
    end subroutine ref_incident_field

    ! Incident_Field_Spherical (Incident_Field.f90:65-166): the CBF excitation
    ! for type_iws = 2 / 3, over the n cells given (an extended block)
    subroutine ref_incident_field_sph(n, x, y, z, ntr, th, ph, t0, t1, E) bind(c, name="ref_incident_field_sph")
        integer(c_int), value :: n, ntr, t0, t1
        real(c_double), intent(in) :: x(n), y(n), z(n), th(ntr), ph(ntr)
        complex(c_double_complex), intent(out) :: E(3*n, 2*(t1-t0+1))
        type(Cell), allocatable :: cc(:)
        type(Dipole), allocatable :: tr(:)
        integer(c_int), allocatable :: id(:)
        real(c_double), allocatable :: zr(:)
        complex(c_double_complex), allocatable :: zc(:)
        integer :: c
        allocate(cc(n), id(n), tr(ntr), zr(n), zc(n))
        id = [(c, c = 1, n)]; zr = 0d0; zc = (0d0, 0d0)
        call build_cells(n, id, x, y, z, zr, zr, zc, zc, cc)
        do c = 1, ntr
            tr(c)%theta = th(c)
            tr(c)%phi = ph(c)
        end do
        call Incident_Field_Spherical(1, n, cc, ntr, tr, t0, t1, E)
    end subroutine ref_incident_field_sph

    ! Green_s_dt (Green_functions.f90:1335-1397): the 3 x 3n dyad from the
    ! cells to a receiver at distance rso in direction (th, ph)
    subroutine ref_green_s_dt(n, x, y, z, kappa, che, rso_in, th, ph, G) bind(c, name="ref_green_s_dt")
        integer(c_int), value :: n
        real(c_double), value :: rso_in, th, ph
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n)
        complex(c_double_complex), intent(out) :: G(3, 3*n)
        type(Cell), allocatable :: cc(:)
        integer(c_int), allocatable :: id(:)
        integer :: c
        allocate(cc(n), id(n))
        id = [(c, c = 1, n)]
        call build_cells(n, id, x, y, z, kappa, kappa, che, che, cc)
        Rso = rso_in
        call Green_s_dt(n, cc, th, ph, G)
    end subroutine ref_green_s_dt

    ! GetFFieldCoeff (Green_functions.f90:1399-1438)
    subroutine ref_ff_coeff(n, x, y, z, kappa, che, th, ph, ff) bind(c, name="ref_ff_coeff")
        integer(c_int), value :: n
        real(c_double), value :: th, ph
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n)
        complex(c_double_complex), intent(out) :: ff(n)
        type(Cell), allocatable :: cc(:)
        integer(c_int), allocatable :: id(:)
        integer :: c
        allocate(cc(n), id(n))
        id = [(c, c = 1, n)]
        call build_cells(n, id, x, y, z, kappa, kappa, che, che, cc)
        call GetFFieldCoeff(n, cc, th, ph, ff)
    end subroutine ref_ff_coeff

    ! Calcul_MatrixZij_ACA (Calcul_ACA_matrices.f90:6-219); curs_* are the
    ! reference's 1-based global row / column cursors.
    subroutine ref_aca(n, x, y, z, kappa, che, curs_lig, curs_col, nrow, ncol, nb_iter, U, V) bind(c, name="ref_aca")
        integer(c_int), value :: n, curs_lig, curs_col, nrow, ncol
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n)
        integer(c_int), intent(out) :: nb_iter
        complex(c_double_complex), intent(out) :: U(nrow, Nb_it_max), V(Nb_it_max, ncol)
        type(Cell), allocatable :: cc(:)
        integer(c_int), allocatable :: id(:)
        integer :: c, k
        allocate(cc(n), id(n))
        id = [(c, c = 1, n)]
        call build_cells(n, id, x, y, z, kappa, kappa, che, che, cc)
        Nbc = n
        call Calcul_MatrixZij_ACA(cc, 1, 2, curs_lig, curs_col, nrow, ncol, k, U, V)
        nb_iter = k
    end subroutine ref_aca

    ! Calcul_GenMatrix_ACA (Calcul_ACA_matrices.f90:222-424)
    subroutine ref_aca_gen(m, n, A, nmax, nout, U, V) bind(c, name="ref_aca_gen")
        integer(c_int), value :: m, n, nmax
        complex(c_double_complex), intent(in) :: A(m, n)
        integer(c_int), intent(out) :: nout
        complex(c_double_complex), intent(out) :: U(m, nmax), V(nmax, n)
        integer :: k
        call Calcul_GenMatrix_ACA(m, n, A, nmax, k, U, V)
        nout = k
    end subroutine ref_aca_gen

    ! Green_s_tr_partial_FN (Green_functions.f90:176-287): Frobenius norm of a block
    subroutine ref_block_fn(n, id, x, y, z, kappa, che, znnpp, fn) bind(c, name="ref_block_fn")
        integer(c_int), value :: n
        integer(c_int), intent(in) :: id(n)
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n), znnpp(n)
        real(c_double), intent(out) :: fn
        type(Cell), allocatable :: cc(:)
        allocate(cc(n))
        call build_cells(n, id, x, y, z, kappa, kappa, che, znnpp, cc)
        call Green_s_tr_partial_FN(n, cc, n, cc, fn)
    end subroutine ref_block_fn

    ! SR_Green_s_tr_partial_FN (Green_functions.f90:564-806): nnz and norm of the SR mask
    subroutine ref_sr_fn(n, id, x, y, z, kappa, che, znnpp, fsr, nnz, fn) bind(c, name="ref_sr_fn")
        integer(c_int), value :: n
        integer(c_int), intent(in) :: id(n)
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n), znnpp(n)
        real(c_double), value :: fsr
        integer(c_int), intent(out) :: nnz
        real(c_double), intent(out) :: fn
        type(Cell), allocatable :: cc(:)
        integer :: k
        allocate(cc(n))
        call build_cells(n, id, x, y, z, kappa, kappa, che, znnpp, cc)
        call SR_Green_s_tr_partial_FN(n, cc, fsr, k, fn)
        nnz = k
    end subroutine ref_sr_fn

    ! SR_Green_s_tr_partial (Green_functions.f90:290-561): the CSR matrix (1-based irow, icol)
    subroutine ref_sr_csr(n, id, x, y, z, kappa, che, znnpp, fsr, nnz, val, irow, icol) bind(c, name="ref_sr_csr")
        integer(c_int), value :: n, nnz
        integer(c_int), intent(in) :: id(n)
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n), znnpp(n)
        real(c_double), value :: fsr
        complex(c_double_complex), intent(out) :: val(nnz)
        integer(c_int), intent(out) :: irow(3 * n + 1), icol(nnz)
        type(Cell), allocatable :: cc(:)
        allocate(cc(n))
        call build_cells(n, id, x, y, z, kappa, kappa, che, znnpp, cc)
        irow = 0
        call SR_Green_s_tr_partial(n, cc, fsr, nnz, val, irow, icol)
    end subroutine ref_sr_csr

    ! setfSR (getParameters_CBFM.f90:113-168): the fSR search of one block
    subroutine ref_setfsr(n, id, x, y, z, kappa, che, znnpp, res_sr_in, fsr, nnz) bind(c, name="ref_setfsr")
        integer(c_int), value :: n
        integer(c_int), intent(in) :: id(n)
        real(c_double), intent(in) :: x(n), y(n), z(n), kappa(n)
        complex(c_double_complex), intent(in) :: che(n), znnpp(n)
        real(c_double), value :: res_sr_in
        real(c_double), intent(out) :: fsr
        integer(c_int), intent(out) :: nnz
        type(Cell), allocatable :: cc(:)
        real(c_double) :: f
        integer :: k
        allocate(cc(n))
        call build_cells(n, id, x, y, z, kappa, kappa, che, znnpp, cc)
        res_SR = res_sr_in
        call setfSR(1, n, cc, 0, f, k)
        fsr = f
        nnz = k
    end subroutine ref_setfsr

    ! getTransmitters_CBFM (getParameters_CBFM.f90:1-110), uniform cos(theta)/phi grid
    subroutine ref_transmitters_un(niws_in, nmax, th, ph, ntr_out) bind(c, name="ref_transmitters_un")
        integer(c_int), value :: niws_in, nmax
        real(c_double), intent(out) :: th(nmax), ph(nmax)
        integer(c_int), intent(out) :: ntr_out
        interface
            subroutine getTransmitters_CBFM(Transmitters_CBFM)
                use Initialization
                type(Dipole), dimension(:), allocatable, intent(out) :: Transmitters_CBFM
            end subroutine getTransmitters_CBFM
        end interface
        type(Dipole), allocatable :: t(:)
        integer :: m
        distr_iws = 1
        Niws = niws_in
        call getTransmitters_CBFM(t)
        ntr_out = size(t)
        m = min(nmax, size(t))
        th(1:m) = t(1:m)%theta
        ph(1:m) = t(1:m)%phi
    end subroutine ref_transmitters_un

    ! getAdaptiveQuadPts2 (Numerical_Integration.f90:213-245): QUADPACK qag with the
    ! reference's 61-point Gauss-Kronrod rule reading the integrand from a grid
    ! (quadpack.f90); tf = 0 for 'th', 1 for 'ph'
    subroutine ref_adaptive_quad(nth, nph, d1, d2, f, tf, jj, a0, b0, res) bind(c, name="ref_adaptive_quad")
        integer(c_int), value :: nth, nph, d1, d2, tf, jj
        real(c_double), intent(in) :: f(d1, d2)
        real(c_double), value :: a0, b0
        real(c_double), intent(out) :: res
        character(len=2) :: t
        external getAdaptiveQuadPts2
        t = 'th'
        if (tf /= 0) t = 'ph'
        res = 0.0d0
        call getAdaptiveQuadPts2(nth, nph, d1, d2, f, t, jj, a0, b0, res)
    end subroutine ref_adaptive_quad

    ! trap_2Dc (Numerical_Integration.f90:251-305)
    subroutine ref_trap2d(nx, ny, f, a, b, c, d, res) bind(c, name="ref_trap2d")
        integer(c_int), value :: nx, ny
        real(c_double), intent(in) :: f(nx, ny)
        real(c_double), value :: a, b, c, d
        real(c_double), intent(out) :: res
        external trap_2Dc
        call trap_2Dc(f, a, b, c, d, res, nx, ny)
    end subroutine ref_trap2d

    ! getGaussLegendreQuadPts3 (Numerical_Integration.f90:88-145, LegendreRuleFast.f90)
    subroutine ref_gauss_legendre(n, a, b, x, w) bind(c, name="ref_gauss_legendre")
        integer(c_int), value :: n
        real(c_double), value :: a, b
        real(c_double), intent(out) :: x(n), w(n)
        external getGaussLegendreQuadPts3
        call getGaussLegendreQuadPts3(n, a, b, x, w)
    end subroutine ref_gauss_legendre

end module midas_ref_driver
