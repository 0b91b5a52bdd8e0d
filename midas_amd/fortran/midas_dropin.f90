! Drop-in replacements for the reference's hot-path subroutines: same names,
! same argument lists and module inputs, computed on the GPU by
! libmidas_gpu.so.  Link this file (with midas_gpu_c.f90) in place of the
! reference's Compute_EFields_MoM.f90, Compute_EFields_CBFME.f90,
! Compute_EFields_CBFME_ACA.f90 and Calcul_ACA_matrices.f90; everything else of
! the reference (input decks, discretisation, block division, getParameters_CBFM,
! output writers) is unchanged.
!
!   Compute_EFields_MoM        Compute_EFields_MoM.f90:1-19
!   Compute_EFields_CBFME      Compute_EFields_CBFME.f90:1-21  (SR mask as the reference
!                              builds it when SR = 1, solved by the dense LU instead of PARDISO)
!   Compute_EFields_CBFME_ACA  Compute_EFields_CBFME_ACA.f90:1-18
!   Calcul_MatrixZij_ACA       Calcul_ACA_matrices.f90:6-20
!   Calcul_GenMatrix_ACA       Calcul_ACA_matrices.f90:222-236
!
! Environment: MIDAS_DEVICE (GPU ordinal, default 0) and MIDAS_SOLVER (fp64,
! c32, f16, f16x3, bf16, bf16x3; default fp64 = the reference's PZGESV
! semantics).  Each MPI rank computes the whole system on its own GPU and
! returns the rows the reference's callers expect of that rank (the CBFM
! outputs: its blocks MPI_CBFM_Blocks(rank+1, 2:), in that order); the intended
! use is one rank per GPU with the problem fitting one GPU's HBM.  A
! mixed-precision solve whose refinement misses its tolerance falls back to the
! FP64 LU and says so.
module midas_dropin_state
    use iso_c_binding
    use midas_gpu_c
    implicit none
    type(c_ptr), save :: midas_context = c_null_ptr
    integer(c_int), save :: midas_mode = MIDAS_SOLVE_FP64
contains
    subroutine midas_dropin_init()
        character(len=32) :: v
        integer :: st, dev
        if (c_associated(midas_context)) return
        dev = 0
        call get_environment_variable("MIDAS_DEVICE", v, status=st)
        if (st == 0) read (v, *) dev
        call get_environment_variable("MIDAS_SOLVER", v, status=st)
        if (st == 0) then
            select case (trim(v))
            case ("fp64"); midas_mode = MIDAS_SOLVE_FP64
            case ("c32"); midas_mode = MIDAS_SOLVE_C32_IR
            case ("f16"); midas_mode = MIDAS_SOLVE_F16_IR
            case ("f16x3"); midas_mode = MIDAS_SOLVE_F16X3_IR
            case ("bf16"); midas_mode = MIDAS_SOLVE_BF16_IR
            case ("bf16x3"); midas_mode = MIDAS_SOLVE_BF16X3_IR
            case default
                write (*, '(a,a)') 'MIDAS GPU: unknown MIDAS_SOLVER ', trim(v)
                error stop 1
            end select
        end if
        if (midas_init(int(dev, c_int), midas_context) /= 0) then
            write (*, '(a,a)') 'MIDAS GPU: ', midas_error_message()
            error stop 1
        end if
    end subroutine midas_dropin_init

    subroutine midas_check(r, what)
        integer(c_int), intent(in) :: r
        character(len=*), intent(in) :: what
        if (r < 0) then
            write (*, '(a,a,a,a)') 'MIDAS GPU: ', what, ': ', midas_error_message()
            error stop 1
        end if
    end subroutine midas_check

    ! report a mixed-precision solve that needed the FP64 fallback
    subroutine midas_report(inf, what)
        type(midas_solve_info_t), intent(in) :: inf
        character(len=*), intent(in) :: what
        if (inf%status == MIDAS_STATUS_FP64_FALLBACK) &
            write (*, '(a,a,a,es10.3,a)') 'MIDAS GPU: ', what, ': refinement stopped at backward error ', &
                inf%berr, '; solved by the FP64 LU instead'
    end subroutine midas_report
end module midas_dropin_state


! The SoA arrays of the cell fields the hot path reads (type Cell is not
! bind(c)); the constants SetCellsParams computed are passed as they are.
module midas_dropin_cells
    use iso_c_binding
    use iso_fortran_env
    use midas_gpu_c
    use Initialization
    implicit none
    type midas_cell_arrays
        real(c_double), allocatable :: x(:), y(:), z(:), sc(:), kappa(:)
        complex(c_double_complex), allocatable :: eps(:), che(:), znnpp(:)
        integer(c_int), allocatable :: id(:)
    end type midas_cell_arrays
contains
    subroutine midas_pack_cells(cells, n, a, cs)
        integer, intent(in) :: n
        type(Cell), intent(in) :: cells(n)
        type(midas_cell_arrays), target, intent(inout) :: a
        type(midas_cells_t), intent(out) :: cs
        a%x = cells%Xc
        a%y = cells%Yc
        a%z = cells%Zc
        a%sc = cells%Sc
        a%kappa = cells%kappa_n
        a%eps = cells%Eps_n
        a%che = cells%Che_n
        a%znnpp = cells%Znnpp
        a%id = cells%n_cell
        cs = midas_cells_t(int(n, c_int), c_loc(a%x), c_loc(a%y), c_loc(a%z), c_loc(a%sc), c_loc(a%eps), &
                           c_loc(a%kappa), c_loc(a%che), c_loc(a%znnpp), c_loc(a%id))
    end subroutine midas_pack_cells
end module midas_dropin_cells


SUBROUTINE Compute_EFields_MoM(Cells,Transmitters,Receivers,S_total,C_ext,C_abs)
    USE Initialization
    USE common_variables
    USE midas_dropin_state
    USE midas_dropin_cells
    IMPLICIT NONE
    type (Cell), Dimension(Nbc), INTENT(IN) :: Cells
    type (Dipole), Dimension(NTr), INTENT(IN) :: Transmitters
    type (Dipole), Dimension(NRx_tot), INTENT(IN) :: Receivers
    COMPLEX(real64), Dimension(NRx_tot,4*NTr), INTENT(OUT), TARGET :: S_total
    COMPLEX(real64), Dimension(NTr), INTENT(OUT), TARGET :: C_ext, C_abs

    type(midas_cell_arrays), target :: ca
    type(midas_cells_t) :: cs
    type(midas_solve_info_t) :: inf
    real(c_double), allocatable :: tth(:), tph(:), rth(:), rph(:)
    integer(c_int) :: r

    call midas_dropin_init()
    call midas_pack_cells(Cells, Nbc, ca, cs)
    tth = Transmitters%theta
    tph = Transmitters%phi
    rth = Receivers%theta
    rph = Receivers%phi
    r = midas_compute_efields_mom(midas_context, cs, k_0, MIDAS_Q_ALL, int(NTr, c_int), tth, tph, &
                                  int(NRx_tot, c_int), rth, rph, midas_mode, c_null_ptr, c_loc(S_total), &
                                  c_loc(C_ext), c_loc(C_abs), c_null_ptr, inf)
    call midas_check(r, 'Compute_EFields_MoM')
    if (rank == 0) call midas_report(inf, 'Compute_EFields_MoM')
    ! the reference prints INFO / "SINGULAR MATRIX" and continues (Compute_EFields_MoM.f90:174-181)
    if (r > 0 .and. rank == 0) write (*, '(a,i10)') ' SINGULAR MATRIX, INFO = ', r
END SUBROUTINE Compute_EFields_MoM


SUBROUTINE midas_cbfm_common(Cells,CBFM_Blocks,CBFM_Blocks_Ext,MPI_CBFM_Blocks,Transmitters,aca_on,threshold, &
                             E_total,K_b,C_b)
    USE Initialization
    USE common_variables
    USE midas_dropin_state
    USE midas_dropin_cells
    IMPLICIT NONE
    type (Cell), Dimension(Nbc), INTENT(IN) :: Cells
    type (CBFM_Block), Dimension(Nblocks), INTENT(IN) :: CBFM_Blocks
    Integer, Dimension(Nblocks,Nbc_ext), INTENT(IN) :: CBFM_Blocks_Ext
    Integer, Dimension(nber_procs,Nblk_proc_max+1), INTENT(IN):: MPI_CBFM_Blocks
    type (Dipole), Dimension(NTr), INTENT(IN) :: Transmitters
    integer, intent(in) :: aca_on
    real(real64), intent(in) :: threshold
    COMPLEX(real64), Dimension(:,:), allocatable, INTENT(OUT), TARGET :: E_total
    Integer, Dimension(:), allocatable, INTENT(OUT) :: K_b
    COMPLEX(real64), Dimension(:,:), allocatable, INTENT(OUT), TARGET :: C_b

    interface
        SUBROUTINE getTransmitters_CBFM(Transmitters_CBFM)
            USE Initialization
            type (Dipole), Dimension(:), allocatable, INTENT(OUT) :: Transmitters_CBFM
        END SUBROUTINE getTransmitters_CBFM
    end interface

    type(midas_cell_arrays), target :: ca
    type(midas_cells_t) :: cs
    type(midas_blocks_t) :: bl
    type(midas_solve_info_t) :: inf
    type (Dipole), Dimension(:), allocatable :: tcbf
    integer(c_int), allocatable, target :: bstart(:), estart(:), ecells(:), kk(:)
    real(c_double), allocatable :: thc(:), phc(:), tth(:), tph(:)
    COMPLEX(real64), Dimension(:,:), allocatable, TARGET :: E_all, C_all
    integer :: b, i, kcap, job, r0, nrow, row
    integer(c_int) :: r, iws

    call midas_dropin_init()
    call midas_pack_cells(Cells, Nbc, ca, cs)
    call getTransmitters_CBFM(tcbf)          ! the reference's own CBF excitation design
    thc = tcbf%theta
    phc = tcbf%phi
    tth = Transmitters%theta
    tph = Transmitters%phi
    allocate(bstart(Nblocks + 1), estart(Nblocks + 1), kk(Nblocks))
    bstart(1) = 0
    estart(1) = 0
    do b = 1, Nblocks
        bstart(b + 1) = bstart(b) + CBFM_Blocks(b)%Nbc_b
        estart(b + 1) = estart(b) + CBFM_Blocks(b)%Nbc_ext
    end do
    allocate(ecells(max(estart(Nblocks + 1), 1)))
    do b = 1, Nblocks
        do i = 1, CBFM_Blocks(b)%Nbc_ext
            ecells(estart(b) + i) = CBFM_Blocks_Ext(b, i) - 1
        end do
    end do
    bl = midas_blocks_t(int(Nblocks, c_int), c_loc(bstart), c_loc(estart), c_loc(ecells))
    kcap = 2 * size(tcbf)
    allocate(E_all(3 * Nbc, 2 * NTr), C_all(3 * Nbc, kcap))
    ! CBF excitation: type_iws in the dense variant (Compute_EFields_CBFME.f90:363-369);
    ! the ACA variant always uses plane waves (Compute_EFields_CBFME_ACA.f90:312)
    iws = int(type_iws, c_int)
    if (aca_on /= 0) iws = 1
    r = midas_compute_efields_cbfm(midas_context, cs, bl, k_0, MIDAS_Q_ALL, int(size(tcbf), c_int), thc, phc, &
                                   iws, int(NTr, c_int), tth, tph, int(aca_on, c_int), int(Nb_it_max, c_int), &
                                   Epsilon_ACA, threshold, midas_mode, c_loc(E_all), int(3 * Nbc, c_long), kk, &
                                   c_loc(C_all), int(3 * Nbc, c_long), int(kcap, c_int), int(SR, c_int), res_SR, &
                                   c_null_ptr, c_null_ptr, inf)
    call midas_check(r, 'Compute_EFields_CBFME')
    if (rank == 0) call midas_report(inf, 'Compute_EFields_CBFME')
    if (r > 0 .and. rank == 0) write (*, '(a,i10)') ' SINGULAR MATRIX, INFO = ', r
    ! this rank's rows: its blocks MPI_CBFM_Blocks(rank+1, 2:) in that order, as
    ! E_total(3*Nbc_proc, 2*NTr), K_patchs(MyNBlocks) and C_job_patchs(3*Nbc_proc, :)
    ! (Compute_EFields_CBFME.f90:154-163, 443-444, 990-1006)
    allocate(E_total(3 * Nbc_proc, 2 * NTr), C_b(3 * Nbc_proc, kcap), K_b(MPI_CBFM_Blocks(rank + 1, 1)))
    C_b = (0d0, 0d0)
    row = 1
    do job = 1, MPI_CBFM_Blocks(rank + 1, 1)
        b = MPI_CBFM_Blocks(rank + 1, 1 + job)
        r0 = 3 * bstart(b) + 1
        nrow = 3 * CBFM_Blocks(b)%Nbc_b
        E_total(row:row + nrow - 1, :) = E_all(r0:r0 + nrow - 1, :)
        C_b(row:row + nrow - 1, :) = C_all(r0:r0 + nrow - 1, :)
        K_b(job) = kk(b)
        row = row + nrow
    end do
    if (row /= 3 * Nbc_proc + 1) then
        write (*, '(a)') 'MIDAS GPU: MPI_CBFM_Blocks of this rank do not hold Nbc_proc cells'
        error stop 1
    end if
END SUBROUTINE midas_cbfm_common


SUBROUTINE Compute_EFields_CBFME(Cells,CBFM_Blocks,CBFM_Blocks_Ext,MPI_CBFM_Blocks,Transmitters,K_patchs, &
                                 C_job_patchs,E_total)
    USE Initialization
    USE common_variables
    IMPLICIT NONE
    type (Cell), Dimension(Nbc), INTENT(IN) :: Cells
    type (CBFM_Block), Dimension(Nblocks), INTENT(IN) :: CBFM_Blocks
    Integer, Dimension(Nblocks,Nbc_ext), INTENT(IN) :: CBFM_Blocks_Ext
    Integer, Dimension(nber_procs,Nblk_proc_max+1), INTENT(IN):: MPI_CBFM_Blocks
    type (Dipole), Dimension(NTr), INTENT(IN) :: Transmitters
    Integer, Dimension(:), allocatable, INTENT(INOUT) :: K_patchs
    COMPLEX(real64), Dimension(:,:), allocatable, INTENT(INOUT) :: C_job_patchs
    COMPLEX(real64), Dimension(:,:), allocatable, INTENT(OUT):: E_total

    interface
        SUBROUTINE midas_cbfm_common(Cells,CBFM_Blocks,CBFM_Blocks_Ext,MPI_CBFM_Blocks,Transmitters,aca_on,threshold, &
                             E_total,K_b,C_b)
            USE Initialization
            USE common_variables
            type (Cell), Dimension(Nbc), INTENT(IN) :: Cells
            type (CBFM_Block), Dimension(Nblocks), INTENT(IN) :: CBFM_Blocks
            Integer, Dimension(Nblocks,Nbc_ext), INTENT(IN) :: CBFM_Blocks_Ext
            Integer, Dimension(nber_procs,Nblk_proc_max+1), INTENT(IN):: MPI_CBFM_Blocks
            type (Dipole), Dimension(NTr), INTENT(IN) :: Transmitters
            integer, intent(in) :: aca_on
            real(real64), intent(in) :: threshold
            COMPLEX(real64), Dimension(:,:), allocatable, INTENT(OUT), TARGET :: E_total
            Integer, Dimension(:), allocatable, INTENT(OUT) :: K_b
            COMPLEX(real64), Dimension(:,:), allocatable, INTENT(OUT), TARGET :: C_b
        END SUBROUTINE midas_cbfm_common
    end interface

    Real(kind=8) :: Threshold_CBFM, m_max_relative
    Real(kind=8), allocatable :: vals(:)
    COMPLEX(real64), Dimension(:,:), allocatable :: C_b
    Integer, Dimension(:), allocatable :: K_b

    ! the reference's threshold rule, evaluated on the same fields (Compute_EFields_CBFME.f90:286-288)
    vals = abs(Cells(1:Nbc)%m_n)
    m_max_relative = ceiling(maxval(vals)) - 2.0
    Threshold_CBFM = max(10.0**(-3.0-m_max_relative),1e-10)
    call midas_cbfm_common(Cells, CBFM_Blocks, CBFM_Blocks_Ext, MPI_CBFM_Blocks, Transmitters, 0, Threshold_CBFM, &
                           E_total, K_b, C_b)
    if (allocated(K_patchs)) deallocate(K_patchs)
    if (allocated(C_job_patchs)) deallocate(C_job_patchs)
    K_patchs = K_b
    C_job_patchs = C_b(:, 1:max(maxval(K_b), 1))
END SUBROUTINE Compute_EFields_CBFME


SUBROUTINE Compute_EFields_CBFME_ACA(Cells,CBFM_Blocks,CBFM_Blocks_Ext,MPI_CBFM_Blocks,Transmitters,E_total)
    USE Initialization
    USE common_variables
    IMPLICIT NONE
    type (Cell), Dimension(Nbc), INTENT(IN) :: Cells
    type (CBFM_Block), Dimension(Nblocks), INTENT(IN) :: CBFM_Blocks
    Integer, Dimension(Nblocks,Nbc_ext), INTENT(IN) :: CBFM_Blocks_Ext
    Integer, Dimension(nber_procs,Nblk_proc_max+1), INTENT(IN):: MPI_CBFM_Blocks
    type (Dipole), Dimension(NTr), INTENT(IN) :: Transmitters
    COMPLEX(real64), Dimension(:,:), allocatable, INTENT(OUT):: E_total

    interface
        SUBROUTINE midas_cbfm_common(Cells,CBFM_Blocks,CBFM_Blocks_Ext,MPI_CBFM_Blocks,Transmitters,aca_on,threshold, &
                             E_total,K_b,C_b)
            USE Initialization
            USE common_variables
            type (Cell), Dimension(Nbc), INTENT(IN) :: Cells
            type (CBFM_Block), Dimension(Nblocks), INTENT(IN) :: CBFM_Blocks
            Integer, Dimension(Nblocks,Nbc_ext), INTENT(IN) :: CBFM_Blocks_Ext
            Integer, Dimension(nber_procs,Nblk_proc_max+1), INTENT(IN):: MPI_CBFM_Blocks
            type (Dipole), Dimension(NTr), INTENT(IN) :: Transmitters
            integer, intent(in) :: aca_on
            real(real64), intent(in) :: threshold
            COMPLEX(real64), Dimension(:,:), allocatable, INTENT(OUT), TARGET :: E_total
            Integer, Dimension(:), allocatable, INTENT(OUT) :: K_b
            COMPLEX(real64), Dimension(:,:), allocatable, INTENT(OUT), TARGET :: C_b
        END SUBROUTINE midas_cbfm_common
    end interface

    Real(kind=8) :: Threshold_CBFM
    COMPLEX(real64), Dimension(:,:), allocatable :: C_b
    Integer, Dimension(:), allocatable :: K_b

    Threshold_CBFM = 1e-3      ! Compute_EFields_CBFME_ACA.f90:124 (a single literal)
    call midas_cbfm_common(Cells, CBFM_Blocks, CBFM_Blocks_Ext, MPI_CBFM_Blocks, Transmitters, 1, Threshold_CBFM, &
                           E_total, K_b, C_b)
END SUBROUTINE Compute_EFields_CBFME_ACA


SUBROUTINE Calcul_MatrixZij_ACA(Cells,I_Z,J_Z,curs_lig_Iz,curs_col_Jz,NbreLigMatZ,NbreColMatZ,nb_iter_out, &
                                Matrix_U,Matrix_V)
    USE Initialization
    USE common_variables
    USE midas_dropin_state
    USE midas_dropin_cells
    IMPLICIT NONE
    type (Cell), Dimension(Nbc), INTENT(IN) :: Cells
    INTEGER, INTENT(IN) :: I_Z,J_Z,curs_lig_Iz,curs_col_Jz,NbreLigMatZ,NbreColMatZ
    INTEGER, INTENT(OUT) :: nb_iter_out
    COMPLEX(real64), Dimension(NbreLigMatZ,Nb_it_max), INTENT(OUT), TARGET :: Matrix_U
    COMPLEX(real64), Dimension(Nb_it_max,NbreColMatZ), INTENT(OUT), TARGET :: Matrix_V

    type(midas_cell_arrays), target :: ca
    type(midas_cells_t) :: cs
    integer(c_int) :: r, k

    call midas_dropin_init()
    call midas_pack_cells(Cells, Nbc, ca, cs)
    r = midas_aca_block(midas_context, cs, k_0, MIDAS_Q_ALL, int((curs_lig_Iz - 1) / 3, c_int), &
                        int(NbreLigMatZ / 3, c_int), int((curs_col_Jz - 1) / 3, c_int), int(NbreColMatZ / 3, c_int), &
                        int(Nb_it_max, c_int), Epsilon_ACA, c_loc(Matrix_U), c_loc(Matrix_V), k)
    call midas_check(r, 'Calcul_MatrixZij_ACA')
    nb_iter_out = k
END SUBROUTINE Calcul_MatrixZij_ACA


SUBROUTINE Calcul_GenMatrix_ACA(m,n,Zin,nmax,nout,Matrix_U,Matrix_V)
    USE Initialization
    USE common_variables
    USE midas_dropin_state
    IMPLICIT NONE
    INTEGER, INTENT(IN) :: m,n,nmax
    COMPLEX(real64), Dimension(m,n), INTENT(IN), TARGET :: Zin
    INTEGER, INTENT(OUT) :: nout
    COMPLEX(real64), Dimension(m,nmax), INTENT(OUT), TARGET :: Matrix_U
    COMPLEX(real64), Dimension(nmax,n), INTENT(OUT), TARGET :: Matrix_V
    integer(c_int) :: r, k

    call midas_dropin_init()
    r = midas_aca_matrix(midas_context, int(m, c_int), int(n, c_int), c_loc(Zin), int(m, c_long), &
                         int(nmax, c_int), Epsilon_ACA, c_loc(Matrix_U), c_loc(Matrix_V), k)
    call midas_check(r, 'Calcul_GenMatrix_ACA')
    nout = k
END SUBROUTINE Calcul_GenMatrix_ACA
