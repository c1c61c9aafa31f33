! Test-infrastructure driver for the reference build in oracle/_ref/.
!
! It replaces only the reference's main program (src/main.f90:4-94) and the
! step loop of spindynamics_run (src/dynamics/spindynamics.f90:260-323); every
! numerical routine it calls is the reference's own:
!   setup_simu                     src/setup/setup-simu.f90:10
!   hamiltonian%get_eff_field      src/Hamiltonian_type/hamiltonian_collection.f90:433
!   hamiltonian%energy             src/Hamiltonian_type/hamiltonian_collection.f90:390
!   select_propagator / get_*_field src/solvers/solver_commun.f90:267
!   get_dt_mode / get_Dmode_int    src/solvers/solver_order.f90:385,417
!   truncate                       src/vector_utils/precision.f90:39
!   init_update_time / update_time src/solvers/update_time.f90:20,41 (called as at
!                                  spindynamics.f90:221,438; a no-op unless the
!                                  input sets `update_time T`)
! The loop is restated (not linked from spindynamics_run) because that routine
! statically reaches the excitation shapes, which need GSL (absent here).
! Differences from spindynamics_run: no EM.dat/topology/IO side channels, and
! the standard normals of the thermal field are read from normals.bin (3N
! doubles per step) instead of being drawn by random_numbers%get_list, so that
! a trajectory at finite T is reproducible.  Everything else of the thermal
! field is the reference's: the generator object of get_ran_type
! (random_public.f90:16-50), its sigma set from the reference's own variables
! as spindynamics.f90:269 sets it, x = mean + sigma*n as rand_normal_int forms
! a draw (random_lib.f90:152-166), and get_temperature_field (langevin_bath).
!
! Usage (in a directory holding the reference's `input` file):
!   mjref_driver field -> field.out: E_total, then B_eff(1:3N)      (ES25.17);
!                         eres.out: per term of the resolved Hamiltonian
!                         (energy_resolved, hamiltonian_collection.f90:365-388)
!                         its description, then its energy
!   mjref_driver traj  -> traj.out: per output step j: j, E/Ncell, m(1:3N);
!                         ts.out: the timestep each step used (ES25.17);
!                         sigma.out: the generator's sigma of each step at
!                         finite T (ES25.17)
program mjref_driver
    use m_derived_types
    use m_setup_simu
    use m_H_public
    use m_hamiltonian_collection, only: hamiltonian
    use m_mpi_start_end
    use m_dyna_io, only: dyna_input
    use m_solver_commun, only: get_integrator_field, get_propagator_field, &
                               select_propagator, get_temperature_field
    use m_solver_order, only: get_dt_mode, get_Dmode_int
    use m_precision, only: truncate
    use m_update_time, only: update_time, init_update_time
    use m_random_public, only: get_ran_type
    use m_random_base, only: ranbase
    use m_constants, only: hbar
    implicit none

    type(io_parameter)          :: io_simu
    type(lattice)               :: lat, lat_1, lat_2
    type(simulation_parameters) :: ext_param
    class(t_H),allocatable      :: Ham_res(:), Ham_comb(:)
    type(hamiltonian)           :: H_res, H_comb
    type(dyna_input)            :: io_dyn
    character(len=32)           :: mode
    real(8),allocatable,target  :: B(:), Dmag(:,:,:), Dmag_int(:,:), Dmag_T(:,:), Dmag_T_zero(:,:)
    real(8),allocatable         :: Eres(:)
    real(8),pointer,contiguous  :: Beff_3(:,:), Dmag_3(:,:,:), Dmag_int_3(:,:), Dmag_T_3(:,:)
    class(ranbase),allocatable,target :: random_numbers
    real(8),pointer,contiguous  :: rand_num_3(:,:)
    real(8),allocatable         :: unit_normals(:)
    logical                     :: used(number_different_order_parameters)
    real(8)                     :: E, dt, kt, timestep_int
    integer                     :: i, j, io, iun, its, isg, N_loop, N_cell, dim_mode, i_loop, n3
    logical                     :: have_noise

    call get_command_argument(1, mode)
    call init_MPI(mpi_world)
    call setup_simu(io_simu, lat, ext_param, Ham_res, Ham_comb, H_res, H_comb)
    dim_mode = lat%M%dim_mode
    N_cell = lat%Ncell
    n3 = dim_mode*N_cell
    allocate(B(n3), source=0.0d0)

    select case (trim(mode))
    case ('field')
        call H_comb%get_eff_field(lat, B, 1)
        E = H_comb%energy(lat)
        open(newunit=io, file='field.out', action='write', status='replace')
        write(io, '(ES25.17)') E
        do i = 1, n3
            write(io, '(ES25.17)') B(i)
        enddo
        close(io)
        allocate(Eres(H_res%N_total))
        call H_res%energy_resolved(lat, Eres)
        open(newunit=io, file='eres.out', action='write', status='replace')
        do i = 1, H_res%N_total
            if (i <= H_res%NH_total) then
                write(io, '(A)') trim(H_res%H(i)%desc)
            else
                write(io, '(A)') 'fft'
            endif
            write(io, '(ES25.17)') Eres(i)
        enddo
        close(io)

    case ('traj')
        call io_dyn%read_file()
        call init_update_time('input')
        timestep_int = io_dyn%timestep
        call lat%used_order(used)
        call lat%copy(lat_1)
        call lat%copy(lat_2)
        kt = ext_param%ktini
        call select_propagator(kt, N_loop)
        allocate(Dmag(dim_mode, N_cell, N_loop), Dmag_int(dim_mode, N_cell), source=0.0d0)
        allocate(Dmag_T(dim_mode, N_cell), Dmag_T_zero(3, N_cell*(dim_mode/3)), source=0.0d0)
        Dmag_3(1:3, 1:N_cell*(dim_mode/3), 1:N_loop) => Dmag
        Dmag_int_3(1:3, 1:N_cell*(dim_mode/3)) => Dmag_int
        Dmag_T_3(1:3, 1:N_cell*(dim_mode/3)) => Dmag_T
        Beff_3(1:3, 1:N_cell*(dim_mode/3)) => B
        call lat%copy_val_to(lat_1)
        ! the generator object as spindynamics_run creates it (spindynamics.f90:237-239)
        call get_ran_type(random_numbers)
        call random_numbers%init_base(dim_mode*N_cell)
        rand_num_3(1:3,1:N_cell*(dim_mode/3)) => random_numbers%x
        allocate(unit_normals(dim_mode*N_cell), source=0.0d0)
        inquire(file='normals.bin', exist=have_noise)
        if (have_noise) then
            open(newunit=iun, file='normals.bin', access='stream', form='unformatted', action='read')
            open(newunit=isg, file='sigma.out', action='write', status='replace')
        endif
        open(newunit=io, file='traj.out', action='write', status='replace')
        open(newunit=its, file='ts.out', action='write', status='replace')
        do j = 1, io_dyn%duration
            write(its, '(ES25.17)') timestep_int
            dt = timestep_int/real(N_loop)
            if (kt*io_dyn%damping > 1.0d-10 .and. have_noise) then
                ! spindynamics.f90:269-271; get_list's draws (rangnu%get_list, random_GNU.f90)
                ! replaced by the standard normals of normals.bin
                call random_numbers%set(sigma=sqrt(io_dyn%damping*kT*hbar/dt))
                read(iun) unit_normals
                if (kt > 1.0d-8) random_numbers%is_set = .true.
                random_numbers%x = random_numbers%mean + random_numbers%sigma*unit_normals
                write(isg, '(ES25.17)') random_numbers%sigma
                call get_temperature_field(random_numbers%is_set, kt, io_dyn%damping, lat_1%M%modes_3, &
                                           Dmag_T_3, rand_num_3)
            endif
            do i_loop = 1, N_loop
                dt = get_dt_mode(dt, i_loop)
                call H_comb%get_eff_field(lat_1, B, 1)
                call get_propagator_field(Beff_3, io_dyn%damping, lat_1%M%modes_3, Dmag_3(:,:,i_loop), io_dyn%SOT)
                call get_Dmode_int(Dmag, i_loop, N_loop, Dmag_int)
                if (i_loop == N_loop) then
                    lat_2%M%modes_3 = get_integrator_field(lat_1%M%modes_3, Dmag_int_3, Dmag_T_3, dt)
                else
                    lat_2%M%modes_3 = get_integrator_field(lat_1%M%modes_3, Dmag_int_3, Dmag_T_zero, dt)
                endif
                call lat_2%M%copy_val(lat_1%M)
            enddo
            call lat_2%M%copy_val(lat%M)
            call truncate(lat, used)
            E = H_comb%energy(lat)/real(N_cell, 8)
            if (mod(j, io_dyn%Efreq) == 0) then
                write(io, '(I10)') j
                write(io, '(ES25.17)') E
                do i = 1, n3
                    write(io, '(ES25.17)') lat%M%all_modes(i)
                enddo
            endif
            call update_time(timestep_int, Beff_3, io_dyn%damping)
        enddo
        close(io)
        close(its)
        if (have_noise) then
            close(iun)
            close(isg)
        endif
    case default
        stop 'usage: mjref_driver field|traj'
    end select
end program
