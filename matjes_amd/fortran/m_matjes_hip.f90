! iso_c_binding interface to libmatjes_hip.so (include/matjes_hip.h).
!
! Conventions follow the reference's own C bindings (cuda_H_interface.f90:5-88,
! cuda_fft_interface.f90:1-89): the context is an opaque type(c_ptr), passed
! by reference to create/destroy (mj_ctx**) and by value elsewhere; arrays are
! assumed-size; every function returns an integer status (0 = success).
module m_matjes_hip
    use, intrinsic :: iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: MJ_OK = 0
    integer(c_int), parameter, public :: MJ_EULER = 1, MJ_HEUN = 2, MJ_RK = 4, MJ_SIB = 5, MJ_DEPONDT = 6
    integer(c_int), parameter, public :: MJ_MC_SPHERE = 0, MJ_MC_ISING = 1
    integer(c_int), parameter, public :: MJ_OP_M = 1, MJ_OP_B = 3

    public :: mj_create, mj_destroy, mj_add_hamiltonian_coo, mj_add_hamiltonian_stencil
    public :: mj_set_bfield, mj_set_fft_operator, mj_set_state, mj_get_state
    public :: mj_eff_field, mj_energy, mj_energy_terms, mj_energy_single, mj_mc_sweeps, mj_mc_set_cone, mj_mc_stats, &
              mj_mc_pt_setup, mj_mc_pt_sweeps, mj_mc_pt_state, mj_mc_pt_get_state, mj_term_mult, &
              mj_set_bfield_maps, mj_steps_driven, &
              mj_steps, mj_steps_noise, mj_steps_adaptive, mj_set_tile, mj_set_rk_order, &
              mj_keep_last_field, mj_last_field, mj_steps_temperature
    public :: mj_error_message

    interface
        integer(c_int) function mj_last_error(buf, len) bind(c, name='mj_last_error')
            import :: c_int, c_char
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_int), value :: len
        end function

        integer(c_int) function mj_create(ctx, device, dim_lat, periodic, dim_mode) bind(c, name='mj_create')
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: ctx
            integer(c_int), value :: device
            integer(c_int), intent(in) :: dim_lat(3), periodic(3)
            integer(c_int), value :: dim_mode
        end function

        integer(c_int) function mj_destroy(ctx) bind(c, name='mj_destroy')
            import :: c_int, c_ptr
            type(c_ptr), intent(inout) :: ctx
        end function

        integer(c_int) function mj_add_hamiltonian_coo(ctx, op_l, op_r, nnz, row1, col1, val) &
                bind(c, name='mj_add_hamiltonian_coo')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: op_l, op_r
            integer(c_int64_t), value :: nnz
            integer(c_int), intent(in) :: row1(*), col1(*)
            real(c_double), intent(in) :: val(*)
        end function

        integer(c_int) function mj_add_hamiltonian_stencil(ctx, nofs, ofs3, blocks9) &
                bind(c, name='mj_add_hamiltonian_stencil')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nofs
            integer(c_int), intent(in) :: ofs3(*)
            real(c_double), intent(in) :: blocks9(*)
        end function

        integer(c_int) function mj_set_bfield(ctx, b, n3) bind(c, name='mj_set_bfield')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: b(*)
            integer(c_int64_t), value :: n3
        end function

        integer(c_int) function mj_set_fft_operator(ctx, dim_mode, N_rep, periodic, K_n) &
                bind(c, name='mj_set_fft_operator')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: dim_mode
            integer(c_int), intent(in) :: N_rep(3), periodic(3)
            real(c_double), intent(in) :: K_n(*)
        end function

        integer(c_int) function mj_set_state(ctx, m, n3) bind(c, name='mj_set_state')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: m(*)
            integer(c_int64_t), value :: n3
        end function

        integer(c_int) function mj_get_state(ctx, m, n3) bind(c, name='mj_get_state')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: m(*)
            integer(c_int64_t), value :: n3
        end function

        integer(c_int) function mj_eff_field(ctx, B, n3) bind(c, name='mj_eff_field')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: B(*)
            integer(c_int64_t), value :: n3
        end function

        ! the Beff of print_Beff (spindynamics.f90:298,400): kept by the last
        ! stage of a call's last step
        integer(c_int) function mj_keep_last_field(ctx, on) bind(c, name='mj_keep_last_field')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
        end function

        integer(c_int) function mj_last_field(ctx, B, n3) bind(c, name='mj_last_field')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: B(*)
            integer(c_int64_t), value :: n3
        end function

        integer(c_int) function mj_energy(ctx, E) bind(c, name='mj_energy')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: E
        end function

        integer(c_int) function mj_energy_terms(ctx, E, nterms) bind(c, name='mj_energy_terms')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: E(*)
            integer(c_int), intent(inout) :: nterms
        end function

        integer(c_int) function mj_energy_single(ctx, n, sites1, spins, E) bind(c, name='mj_energy_single')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n
            integer(c_int64_t), intent(in) :: sites1(*)
            type(c_ptr), value :: spins
            real(c_double), intent(out) :: E(*)
        end function

        integer(c_int) function mj_mc_sweeps(ctx, nsweeps, sampler, kT, seed, sweep0, measure_every, &
                                             E_samples, M_samples) bind(c, name='mj_mc_sweeps')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsweeps, sampler
            real(c_double), value :: kT
            integer(c_int64_t), value :: seed, sweep0
            integer(c_int), value :: measure_every
            type(c_ptr), value :: E_samples, M_samples
        end function

        integer(c_int) function mj_mc_set_cone(ctx, cone) bind(c, name='mj_mc_set_cone')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: cone
        end function

        integer(c_int) function mj_mc_stats(ctx, out4, colouring3) bind(c, name='mj_mc_stats')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out4(4)
            type(c_ptr), value :: colouring3
        end function

        integer(c_int) function mj_mc_pt_setup(ctx, nrep, kT_ladder, cone) bind(c, name='mj_mc_pt_setup')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nrep
            real(c_double), intent(in) :: kT_ladder(*)
            real(c_double), value :: cone
        end function

        integer(c_int) function mj_mc_pt_sweeps(ctx, nsweeps, sampler, seed, sweep0, exchange_every, &
                                                measure_every, E_samples, M_samples) bind(c, name='mj_mc_pt_sweeps')
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsweeps, sampler
            integer(c_int64_t), value :: seed, sweep0
            integer(c_int), value :: exchange_every, measure_every
            type(c_ptr), value :: E_samples, M_samples
        end function

        integer(c_int) function mj_mc_pt_state(ctx, image_of_rung, swaps_accepted, swaps_attempted, &
                                               replica_stats) bind(c, name='mj_mc_pt_state')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, image_of_rung, swaps_accepted, swaps_attempted, replica_stats
        end function

        integer(c_int) function mj_mc_pt_get_state(ctx, rung, m, n3) bind(c, name='mj_mc_pt_get_state')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: rung
            real(c_double), intent(out) :: m(*)
            integer(c_int64_t), value :: n3
        end function

        integer(c_int) function mj_set_bfield_maps(ctx, nmaps, maps, n) bind(c, name='mj_set_bfield_maps')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nmaps
            real(c_double), intent(in) :: maps(*)
            integer(c_int64_t), value :: n
        end function

        integer(c_int) function mj_steps_driven(ctx, nsteps, integrator, timestep, damping, kT, seed, step0, &
                                                weights, weights_end, energies) bind(c, name='mj_steps_driven')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps, integrator
            real(c_double), value :: timestep, damping, kT
            integer(c_int64_t), value :: seed, step0
            real(c_double), intent(in) :: weights(*)
            type(c_ptr), value :: weights_end, energies
        end function

        integer(c_int) function mj_term_mult(ctx, term, side, vec, nvec, res, nres, alpha, beta) &
                bind(c, name='mj_term_mult')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: term, side
            real(c_double), intent(in) :: vec(*)
            integer(c_int64_t), value :: nvec
            real(c_double), intent(inout) :: res(*)
            integer(c_int64_t), value :: nres
            real(c_double), value :: alpha, beta
        end function

        integer(c_int) function mj_set_rk_order(ctx, n_order) bind(c, name='mj_set_rk_order')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: n_order
        end function

        integer(c_int) function mj_steps(ctx, nsteps, integrator, timestep, damping, kT, seed, step0, energies) &
                bind(c, name='mj_steps')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps, integrator
            real(c_double), value :: timestep, damping, kT
            integer(c_int64_t), value :: seed, step0
            type(c_ptr), value :: energies
        end function

        ! one kT per step (a temperature schedule such as update_TPulse's)
        integer(c_int) function mj_steps_temperature(ctx, nsteps, integrator, timestep, damping, kT_steps, seed, &
                                                     step0, energies) bind(c, name='mj_steps_temperature')
            import :: c_int, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps, integrator
            real(c_double), value :: timestep, damping
            real(c_double), intent(in) :: kT_steps(*)
            integer(c_int64_t), value :: seed, step0
            type(c_ptr), value :: energies
        end function

        integer(c_int) function mj_steps_noise(ctx, nsteps, integrator, timestep, damping, kT, noise, energies) &
                bind(c, name='mj_steps_noise')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps, integrator
            real(c_double), value :: timestep, damping, kT
            real(c_double), intent(in) :: noise(*)
            type(c_ptr), value :: energies
        end function

        ! update_time T (src/solvers/update_time.f90:41-75): timestep is in/out
        integer(c_int) function mj_steps_adaptive(ctx, nsteps, integrator, timestep, damping, kT, seed, step0, &
                discretization, energies, timesteps) bind(c, name='mj_steps_adaptive')
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps, integrator
            real(c_double), intent(inout) :: timestep
            real(c_double), value :: damping, kT, discretization
            integer(c_int64_t), value :: seed, step0
            type(c_ptr), value :: energies, timesteps
        end function

        integer(c_int) function mj_set_tile(ctx, bx, by, bz) bind(c, name='mj_set_tile')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: bx, by, bz
        end function
    end interface

contains

    function mj_error_message() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char) :: buf(1024)
        integer :: i, n, rc
        rc = mj_last_error(buf, 1024_c_int)
        n = 0
        do i = 1, 1024
            if (buf(i) == c_null_char) exit
            n = i
        enddo
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = buf(i)
        enddo
    end function
end module m_matjes_hip
