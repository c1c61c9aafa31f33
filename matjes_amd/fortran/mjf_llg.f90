! Fortran host driver of the device LLG path (host code in the reference's
! language, calling HIP only through m_matjes_hip / iso_c_binding).
!
! It does what spindynamics_run does with a device Hamiltonian
! (src/dynamics/spindynamics.f90:205-323): register the assembled COO terms
! once (as set_Hamiltonians/combine_Hamiltonians produce them), upload the
! initial state, evaluate B_eff and E (hamiltonian%get_eff_field / energy),
! then advance nsteps with the energy after every step.
!
! Input  case.bin (stream): int32 nx,ny,nz,px,py,pz,integrator,nsteps
!                           real64 timestep,damping,kT ; int64 seed
!                           int64 nnz_mm ; int32 row(nnz),col(nnz) ; real64 val(nnz)
!                           int64 nnz_z  ; int32 row(nnz),col(nnz) ; real64 val(nnz)
!                           real64 b(3N) ; real64 m0(3N)
! Output out.bin   (stream): real64 B(3N), E0, E(nsteps), m(3N)
program mjf_llg
    use, intrinsic :: iso_c_binding
    use m_matjes_hip
    implicit none
    type(c_ptr) :: ctx
    integer(c_int) :: hdr(8), dims(3), per(3), integ, nsteps, rc
    real(c_double) :: prm(3), E0
    integer(c_int64_t) :: seed, nnz, n3
    integer(c_int), allocatable :: row(:), col(:)
    real(c_double), allocatable, target :: val(:), b(:), m(:), B_eff(:), E(:)
    integer :: iu

    open(newunit=iu, file='case.bin', access='stream', form='unformatted', action='read', status='old')
    read(iu) hdr
    dims = hdr(1:3); per = hdr(4:6); integ = hdr(7); nsteps = hdr(8)
    read(iu) prm
    read(iu) seed

    rc = mj_create(ctx, 0_c_int, dims, per, 3_c_int)
    call check(rc, 'mj_create')
    n3 = 3_c_int64_t*int(dims(1), c_int64_t)*int(dims(2), c_int64_t)*int(dims(3), c_int64_t)

    read(iu) nnz
    allocate(row(nnz), col(nnz), val(nnz))
    read(iu) row, col, val
    call check(mj_add_hamiltonian_coo(ctx, MJ_OP_M, MJ_OP_M, nnz, row, col, val), 'M-M term')
    deallocate(row, col, val)
    read(iu) nnz
    allocate(row(nnz), col(nnz), val(nnz))
    read(iu) row, col, val
    if (nnz > 0) call check(mj_add_hamiltonian_coo(ctx, MJ_OP_B, MJ_OP_M, nnz, row, col, val), 'Zeeman term')
    allocate(b(n3), m(n3), B_eff(n3), E(max(nsteps, 1)))
    read(iu) b, m
    close(iu)
    if (nnz > 0) call check(mj_set_bfield(ctx, b, n3), 'mj_set_bfield')

    call check(mj_set_state(ctx, m, n3), 'mj_set_state')
    call check(mj_eff_field(ctx, B_eff, n3), 'mj_eff_field')
    call check(mj_energy(ctx, E0), 'mj_energy')
    call check(mj_steps(ctx, nsteps, integ, prm(1), prm(2), prm(3), seed, 0_c_int64_t, c_loc(E)), 'mj_steps')
    call check(mj_get_state(ctx, m, n3), 'mj_get_state')
    call check(mj_destroy(ctx), 'mj_destroy')

    open(newunit=iu, file='out.bin', access='stream', form='unformatted', action='write', status='replace')
    write(iu) B_eff, E0, E(1:nsteps), m
    close(iu)

contains
    subroutine check(code, what)
        integer(c_int), intent(in) :: code
        character(len=*), intent(in) :: what
        if (code /= MJ_OK) then
            write(*, '(3A,I4,2A)') 'mjf_llg: ', what, ' failed (', code, '): ', mj_error_message()
            error stop 1
        endif
    end subroutine
end program mjf_llg
