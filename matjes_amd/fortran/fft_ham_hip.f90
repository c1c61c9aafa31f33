! FFT Hamiltonian backend on the MI355X library: an fft_H (the reference's
! Hamiltonian_type/fft_ham_base.f90 interface, as its fft_H_fftw and
! fft_H_cufft are) whose transform and k-space product run on the device
! (mj_set_fft_operator, mj_eff_field).  Added to the reference's source list,
! it is selected like the other implementations by get_fft_H_single /
! get_fft_H_N (Hamiltonian_type/fft_ham_public.f90:59-122), and the reference's
! operator builders (get_dipolar_fft, get_exchange_J_fft, ...) fill it through
! init_shape and init_op as they fill those.
!
! Each term owns one library context holding only its operator: created in
! init_shape (which has the lattice shape), given the operator in init_op.
! The real-space operator is kept on the host (as init_op received it, before
! the 1/prod(N_rep) of fft_ham_cufft.f90:222, which the library applies), so
! that copy and add can build the operator of another context.
!
!   get_H: Hout = IFFT(K_F^T . FFT(M)) on the lattice sites
!          (fft_ham_fftw.f90:239-262); the library returns the field
!          -2 Hout (hamiltonian_collection.f90:450-453) of a context that holds
!          only this term, and -0.5 of it is Hout exactly
! get_E, get_E_single, get_E_distrib: inherited (through get_H / get_H_single)
module m_fft_H_hip
    use iso_c_binding
    use, intrinsic :: iso_fortran_env, only: error_unit
    use m_fft_H_base, only: fft_H
    use m_type_lattice, only: lattice
    use m_matjes_hip
    implicit none

    private
    public :: fft_H_hip

    type, extends(fft_H) :: fft_H_hip
        private
        type(c_ptr) :: mj = c_null_ptr
        integer :: dim_lat(3) = 0
        real(8), allocatable :: K_n(:, :, :)
    contains
        procedure, public :: get_H
        procedure, public :: get_H_single
        procedure, public :: init_shape
        procedure, public :: init_op
        procedure, public :: destroy
        procedure, public :: copy
        procedure, public :: mv
        procedure, public :: add
        procedure, public :: bcast => bcast_hip
        procedure :: set_M
    end type

contains

    subroutine hip_fail(what)
        character(len=*), intent(in) :: what
        write (error_unit, '(3A)') what, ': ', mj_error_message()
        error stop
    end subroutine

    ! a context for this shape (dim_lat and periodic as init_shape received them)
    subroutine new_context(this)
        class(fft_H_hip), intent(inout) :: this
        integer(c_int) :: rc, per(3)
        if (this%dim_mode /= 3) error stop "fft_H_hip: one magnetic atom per cell (dim_mode 3) only"
        per = merge(1_c_int, 0_c_int, this%periodic)
        rc = mj_create(this%mj, 0_c_int, int(this%dim_lat, c_int), per, 3_c_int)
        if (rc /= MJ_OK) call hip_fail('fft_H_hip: mj_create')
    end subroutine

    subroutine put_operator(this)
        class(fft_H_hip), intent(inout) :: this
        integer(c_int) :: rc, per(3)
        per = merge(1_c_int, 0_c_int, this%periodic)
        rc = mj_set_fft_operator(this%mj, int(this%dim_mode, c_int), int(this%N_rep, c_int), per, this%K_n)
        if (rc /= MJ_OK) call hip_fail('fft_H_hip: mj_set_fft_operator')
    end subroutine

    subroutine init_shape(this, dim_mode, periodic, dim_lat, Kbd, N_rep)
        class(fft_H_hip), intent(inout) :: this
        integer, intent(in) :: dim_mode
        logical, intent(in) :: periodic(3)
        integer, intent(in) :: dim_lat(3)
        integer, intent(out) :: Kbd(2, 3)
        integer, intent(out) :: N_rep(3)

        if (c_associated(this%mj)) error stop "fft_H_hip: init_shape shall only be called once"
        call this%fft_H%init_shape(dim_mode, periodic, dim_lat, Kbd, N_rep)
        this%dim_lat = dim_lat
        call new_context(this)
    end subroutine

    subroutine init_op(this, dim_mode, K_n, desc_in)
        class(fft_H_hip), intent(inout) :: this
        integer, intent(in) :: dim_mode
        real(8), intent(inout), allocatable :: K_n(:, :, :)
        character(len=*), intent(in), optional :: desc_in

        if (.not. c_associated(this%mj)) error stop "fft_H_hip: init_op before init_shape"
        if (any(shape(K_n) /= [this%dim_mode, this%dim_mode, product(this%N_rep)])) &
            error stop "fft_H_hip: operator shape is not (dim_mode, dim_mode, prod(N_rep))"
        call this%fft_H%init_op(dim_mode, K_n, desc_in)
        call move_alloc(K_n, this%K_n)
        call put_operator(this)
        this%set = .true.
    end subroutine

    subroutine set_M(this, lat)
        class(fft_H_hip), intent(inout) :: this
        type(lattice), intent(in) :: lat
        integer(c_int) :: rc
        rc = mj_set_state(this%mj, lat%M%all_modes, int(size(lat%M%all_modes), c_int64_t))
        if (rc /= MJ_OK) call hip_fail('fft_H_hip: mj_set_state')
    end subroutine

    subroutine get_H(this, lat, Hout)
        class(fft_H_hip), intent(inout) :: this
        type(lattice), intent(in) :: lat
        real(8), intent(inout) :: Hout(:, :)
        real(8), allocatable :: B(:)
        integer(c_int) :: rc

        if (.not. this%set) error stop "fft_H_hip: get_H before init_op"
        call this%set_M(lat)
        allocate (B(size(Hout)))
        rc = mj_eff_field(this%mj, B, int(size(B), c_int64_t))
        if (rc /= MJ_OK) call hip_fail('fft_H_hip: mj_eff_field')
        Hout = reshape(-0.5d0*B, shape(Hout))
    end subroutine

    subroutine get_H_single(this, lat, site, Hout)
        class(fft_H_hip), intent(inout) :: this
        type(lattice), intent(in) :: lat
        integer, intent(in) :: site
        real(8), intent(inout) :: Hout(3)
        real(8), allocatable :: Hall(:, :)

        allocate (Hall(3, lat%Ncell))
        call this%get_H(lat, Hall)
        Hout = Hall(:, site)
    end subroutine

    subroutine copy(this, H_out)
        class(fft_H_hip), intent(in) :: this
        class(fft_H), intent(inout) :: H_out

        call this%fft_H%copy(H_out)
        select type (H_out)
        type is (fft_H_hip)
            H_out%dim_lat = this%dim_lat
            allocate (H_out%K_n, source=this%K_n)
            call new_context(H_out)
            call put_operator(H_out)
        class default
            error stop "fft_H_hip: copy into another fft_H implementation"
        end select
    end subroutine

    ! the sum of two operators of the same space (K_F is linear in K_n)
    subroutine add(this, H_in)
        class(fft_H_hip), intent(inout) :: this
        class(fft_H), intent(in) :: H_in

        if (.not. H_in%is_set()) error stop "fft_H_hip: add of an fft_H that is not set"
        if (.not. this%is_set()) then
            call H_in%copy(this)
            return
        endif
        if (.not. this%same_space(H_in)) error stop "fft_H_hip: add of an fft_H on another space"
        select type (H_in)
        type is (fft_H_hip)
            call this%fft_H%add(H_in)
            this%K_n = this%K_n + H_in%K_n
            call put_operator(this)
        class default
            error stop "fft_H_hip: add of another fft_H implementation"
        end select
    end subroutine

    subroutine mv(this, H_out)
        class(fft_H_hip), intent(inout) :: this
        class(fft_H), intent(inout) :: H_out

        select type (H_out)
        type is (fft_H_hip)
            call this%fft_H%mv(H_out)
            H_out%mj = this%mj
            H_out%dim_lat = this%dim_lat
            call move_alloc(this%K_n, H_out%K_n)
            this%mj = c_null_ptr
        class default
            error stop "fft_H_hip: mv into another fft_H implementation"
        end select
        call this%destroy()
    end subroutine

    subroutine destroy(this)
        class(fft_H_hip), intent(inout) :: this
        integer(c_int) :: rc

        call this%fft_H%destroy()
        if (c_associated(this%mj)) rc = mj_destroy(this%mj)
        this%mj = c_null_ptr
        this%dim_lat = 0
        if (allocated(this%K_n)) deallocate (this%K_n)
    end subroutine

    ! one process per device: the MPI distribution of the reference is not used
    subroutine bcast_hip(this, comm)
        use mpi_basic, only: mpi_type
        class(fft_H_hip), intent(inout) :: this
        type(mpi_type), intent(in) :: comm
        error stop "fft_H_hip: no MPI distribution (one process per device)"
    end subroutine
end module m_fft_H_hip
