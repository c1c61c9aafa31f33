! Hamiltonian backend on the MI355X library: a t_H_coo_based type (the
! reference's Hamiltonian_type/ham_coo_based.f90 interface, as its t_H_dense
! and t_H_cusparse are) whose products run on the device through the C-ABI
! (mj_term_mult).  Added to the reference's source list, it is selected like
! any other backend by get_Htype (Hamiltonian_type/ham_public.f90:21-69).
!
! Each term owns one library context holding just that term.  The context is
! created at the first product (set_from_Hcoo has no lattice; the products
! carry it), from the COO the term was built with; copy and add work on that
! COO and drop the context so that the next product rebuilds it.
!
!   mult_r: res = alpha H . mode_r + beta res     (ham_dense.f90:30-59)
!   mult_l: res = alpha mode_l . H + beta res     (ham_dense.f90:61-87), with the
!           dense backend's treatment of the optional arguments (alpha only adds res)
module m_H_hip
    use iso_c_binding
    use m_derived_types, only: lattice
    use m_H_coo_based
    use m_H_type, only: t_H_base
    use m_work_ham_single, only: work_mode, N_work
    use m_matjes_hip
    use, intrinsic :: iso_fortran_env, only: error_unit
    implicit none

    private
    public :: t_H, t_H_hip

    ! the library context of a term: reached through a pointer component so that
    ! the products (intent(in) on the term in the t_H_base interface) may create it
    type :: hip_box
        type(c_ptr) :: mj = c_null_ptr
    end type

    type, extends(t_H_coo_based) :: t_H_hip
        type(hip_box), pointer :: box => null()
        integer :: nnz = 0
        real(8), allocatable :: val(:)
        integer, allocatable :: rowind(:), colind(:)
    contains
        procedure :: set_from_Hcoo
        procedure :: add_child
        procedure :: destroy_child
        procedure :: copy_child
        procedure :: optimize
        procedure :: mult_r, mult_l
        procedure :: send
        procedure :: recv
        procedure :: distribute
        procedure :: bcast
    end type

contains

    subroutine hip_fail(what)
        character(len=*), intent(in) :: what
        write (error_unit, '(3A)') what, ': ', mj_error_message()
        error stop
    end subroutine

    ! the term's own context, from its COO (1-based rows/columns of the reference)
    subroutine product(this, lat, side, vec, res, alpha, beta)
        class(t_H_hip), intent(in) :: this
        type(lattice), intent(in) :: lat
        integer, intent(in) :: side
        real(8), intent(in) :: vec(:)
        real(8), intent(inout) :: res(:)
        real(8), intent(in) :: alpha, beta
        integer(c_int) :: rc, dims(3), per(3)

        if (.not. associated(this%box)) stop "t_H_hip: term not set"
        if (.not. c_associated(this%box%mj)) then
            dims = int(lat%dim_lat, c_int)
            per = merge(1_c_int, 0_c_int, lat%periodic)
            rc = mj_create(this%box%mj, 0_c_int, dims, per, int(this%dim_mode(2), c_int))
            if (rc /= MJ_OK) call hip_fail('t_H_hip: mj_create')
            rc = mj_add_hamiltonian_coo(this%box%mj, int(this%op_l(1), c_int), int(this%op_r(1), c_int), &
                                        int(this%nnz, c_int64_t), this%rowind, this%colind, this%val)
            if (rc /= MJ_OK) call hip_fail('t_H_hip: mj_add_hamiltonian_coo')
        endif
        rc = mj_term_mult(this%box%mj, 0_c_int, int(side, c_int), vec, int(size(vec), c_int64_t), res, &
                          int(size(res), c_int64_t), alpha, beta)
        if (rc /= MJ_OK) call hip_fail('t_H_hip: mj_term_mult')
    end subroutine

    subroutine drop_context(this)
        class(t_H_hip), intent(inout) :: this
        integer(c_int) :: rc
        if (.not. associated(this%box)) allocate (this%box)
        if (c_associated(this%box%mj)) rc = mj_destroy(this%box%mj)
        this%box%mj = c_null_ptr
    end subroutine

    subroutine mult_r(this, lat, res, work, alpha, beta)
        class(t_H_hip), intent(in) :: this
        type(lattice), intent(in) :: lat
        real(8), intent(inout) :: res(:)
        type(work_mode), intent(inout) :: work
        real(8), intent(in), optional :: alpha
        real(8), intent(in), optional :: beta
        real(8), pointer, contiguous :: modes(:)
        integer :: work_size(N_work)
        real(8) :: a, b

        if (size(res) /= this%dimH(1)) stop "size of vec is wrong"
        call this%mode_r%get_mode(lat, modes, work, work_size)
        a = 1.0d0
        b = 0.0d0
        if (present(alpha)) a = alpha
        if (present(beta)) b = beta
        call product(this, lat, 0, modes, res, a, b)
        nullify (modes)
        work%offset = work%offset - work_size
    end subroutine

    subroutine mult_l(this, lat, res, work, alpha, beta)
        class(t_H_hip), intent(in) :: this
        type(lattice), intent(in) :: lat
        real(8), intent(inout) :: res(:)
        type(work_mode), intent(inout) :: work
        real(8), intent(in), optional :: alpha
        real(8), intent(in), optional :: beta
        real(8), pointer, contiguous :: modes(:)
        integer :: work_size(N_work)
        real(8) :: a, b

        if (size(res) /= this%dimH(2)) stop "size of vec is wrong"
        call this%mode_l%get_mode(lat, modes, work, work_size)
        a = 1.0d0
        b = 0.0d0
        if (present(alpha)) a = alpha
        if (present(beta)) b = beta
        if (present(alpha) .and. .not. present(beta)) b = 1.0d0
        call product(this, lat, 1, modes, res, a, b)
        nullify (modes)
        work%offset = work%offset - work_size
    end subroutine

    subroutine optimize(this)
        class(t_H_hip), intent(inout) :: this
        continue
    end subroutine

    subroutine copy_child(this, Hout)
        class(t_H_hip), intent(in) :: this
        class(t_H_base), intent(inout) :: Hout
        select type (Hout)
        class is (t_H_hip)
            Hout%nnz = this%nnz
            allocate (Hout%val, source=this%val)
            allocate (Hout%rowind, source=this%rowind)
            allocate (Hout%colind, source=this%colind)
            Hout%box => null()
            allocate (Hout%box)
            call this%copy_deriv(Hout)
        class default
            stop "Cannot copy t_H_hip type with Hamiltonian that is not a class of t_H_hip"
        end select
    end subroutine

    subroutine add_child(this, H_in)
        class(t_H_hip), intent(inout) :: this
        class(t_H_base), intent(in) :: H_in
        select type (H_in)
        class is (t_H_hip)
            this%val = [this%val, H_in%val]
            this%rowind = [this%rowind, H_in%rowind]
            this%colind = [this%colind, H_in%colind]
            this%nnz = this%nnz + H_in%nnz
            call drop_context(this)
        class default
            stop "Cannot add t_H_hip type with Hamiltonian that is not a class of t_H_hip"
        end select
    end subroutine

    subroutine destroy_child(this)
        class(t_H_hip), intent(inout) :: this
        call drop_context(this)
        deallocate (this%box)
        if (allocated(this%val)) deallocate (this%val)
        if (allocated(this%rowind)) deallocate (this%rowind)
        if (allocated(this%colind)) deallocate (this%colind)
        this%nnz = 0
    end subroutine

    subroutine set_from_Hcoo(this, H_coo)
        class(t_H_hip), intent(inout) :: this
        type(t_H_coo), intent(inout) :: H_coo
        call drop_context(this)
        call H_coo%pop_par(this%dimH, this%nnz, this%val, this%rowind, this%colind)
    end subroutine

    ! one process per device: the MPI distribution of the reference is not used
    subroutine send(this, ithread, tag, com)
        class(t_H_hip), intent(in) :: this
        integer, intent(in) :: ithread, tag, com
        continue
    end subroutine

    subroutine recv(this, ithread, tag, com)
        class(t_H_hip), intent(inout) :: this
        integer, intent(in) :: ithread, tag, com
        continue
    end subroutine

    subroutine bcast(this, comm)
        use mpi_basic, only: mpi_type
        class(t_H_hip), intent(inout) :: this
        type(mpi_type), intent(in) :: comm
        continue
    end subroutine

    subroutine distribute(this, comm)
        use mpi_basic, only: mpi_type
        class(t_H_hip), intent(inout) :: this
        type(mpi_type), intent(in) :: comm
        error stop "t_H_hip: no inner MPI distribution (one process per device)"
    end subroutine
end module m_H_hip
