!*******************************************************************************
!< Driver of the GPU Parker pusher: the reference program's command line
!< (stochastic-mhd.f90:572-1107, same short and long flag names and defaults)
!< and its time loop (solve_transport_equation, stochastic-mhd.f90:312-567),
!< calling the gpat_hip module procedures that stand in for the reference's
!< particle_module / mhd_data_parallel / diagnostics procedures.
!<
!< GPU-only flags: -gp/--gpu_precision (f64 = reference arithmetic, default;
!< f32 = FP32 gather/coefficients with FP64 positions and time; f32h = f32 on
!< records with FP16 gradients),
!< -rs/--rng_seed (Philox key) and -ss/--sort_steps (spatial re-sort of the
!< particles every that many steps inside an interval; 0 = once per interval).
!*******************************************************************************
program stochastic_hip
    use iso_c_binding
    use gpat_hip
    implicit none
    integer, parameter :: dp = kind(1.0d0)
    character(len=256) :: dir_mhd_data = '', diagnostics_directory = 'data/'
    character(len=64) :: conf_file = 'conf.dat', mhd_config_filename = 'mhd_config.dat'
    character(len=8) :: gpu_precision = 'f64'
    integer(c_int64_t) :: nptl_max = 10000000_c_int64_t, rng_seed = 20261202_c_int64_t
    integer :: nptl = 10000
    real(dp) :: dt = 1.0d-7, power_index = 7.0_dp, split_ratio = 2.72_dp, pmin_split = 2.0_dp
    real(dp) :: particle_v0 = 5.0_dp, kperp_kpara = 0.01_dp, quota_hour = 48.0_dp
    real(dp) :: drift_param1 = 4.0d7, drift_param2 = 2.0d8
    real(dp) :: ptl_xmin = 0.0_dp, ptl_xmax = 1.0_dp, ptl_ymin = 0.0_dp, ptl_ymax = 1.0_dp
    real(dp) :: ptl_zmin = 0.0_dp, ptl_zmax = 1.0_dp
    integer :: t_start = 0, t_end = 200, tmax_mhd = 100000, time_interp_flag = 0
    integer :: single_time_frame = 0, dist_flag = 0, split_flag = 1, nsteps_interval = 10
    integer :: num_fine_steps = 1, tmax_to_inject = 100000, ndim_field = 2, charge = -1
    integer :: sort_steps = 0
    integer :: uniform_grid = 0, spherical_coord = 0, check_drift_2d = 0
    integer :: dpp_wave = 0, dpp_shear = 0, weak_scattering = 0
    real(dp) :: tau0_scattering = 1.0_dp
    integer :: deltab_flag = 0, correlation_flag = 0
    integer :: include_3rd_dim = 0, acc_by_surface = 0, size_mpi_sub = 1
    logical :: focused_transport = .false., nlgc = .true., local_dist = .false.
    logical :: inject_new_ptl = .false., inject_part_box = .false., restart_flag = .false.
    logical :: dump_escaped_dist = .false., track_particle_flag = .false.
    logical :: inject_at_shock = .false., inject_large_jz = .false., varying_dt_mhd = .false.
    logical :: inject_same_nptl = .false., inject_large_absj = .false., inject_large_db2 = .false.
    logical :: inject_large_divv = .false., inject_large_rho = .false.
    real(dp) :: jz_min = 100.0_dp, absj_min = 5.0_dp, divv_min = 10.0_dp, rho_min = 5.0_dp
    integer :: ncells_large_jz_norm = 800, ncells_large_absj_norm = 800
    integer :: ncells_large_divv_norm = 800, ncells_large_rho_norm = 800
    real(dp) :: part_box(6)
    real(dp) :: step1, step2, start, finish
    integer :: tf, tmin
    character(len=256) :: fname

    call cpu_time(start)
    call get_cmd_args
    if (single_time_frame == 1) time_interp_flag = 0
    call check_scope

    call load_mhd_config(trim(dir_mhd_data)//trim(mhd_config_filename))
    call init_tstamps_mhd(t_start, t_end)
    if (varying_dt_mhd) then
        call load_tstamps_mhd(t_start, t_end, tmax_mhd, dir_mhd_data)
    else
        call calc_tstamps_mhd(t_start, t_end)
    endif
    call init_field_data(time_interp_flag, ndim_field)
    call read_particle_boundary_conditions('config/'//trim(conf_file))
    call read_particle_params('config/'//trim(conf_file))
    call read_diagnostics_params('config/'//trim(conf_file), t_end - t_start, local_dist)
    call set_drift_parameters(drift_param1, drift_param2, charge)
    if (ndim_field == 2 .and. check_drift_2d == 1) call set_flag_check_drift_2d(check_drift_2d)
    call set_dpp_params(dpp_wave, dpp_shear, weak_scattering, tau0_scattering)
    call set_nlgc(nlgc, kperp_kpara)
    select case (trim(gpu_precision))
    case ('f32'); call set_precision(1, rng_seed)
    case ('f32h'); call set_precision(2, rng_seed)
    case default; call set_precision(0, rng_seed)
    end select
    call set_sort_steps(sort_steps)
    call init_particles(nptl_max)
    call init_particle_distributions(local_dist, dump_escaped_dist)
    tmin = t_start

    ! solve_transport_equation
    write(fname, "(A,I4.4)") trim(dir_mhd_data)//'mhd_data_', tmin
    call read_field_data_parallel(fname, 0)
    call calc_fields_gradients(0)
    if (inject_part_box) then
        part_box = [ptl_xmin, ptl_ymin, ptl_zmin, ptl_xmax, ptl_ymax, ptl_zmax]
    else
        part_box = [mhd_config%xmin, mhd_config%ymin, mhd_config%zmin, &
                    mhd_config%xmin + mhd_config%nx * mhd_config%dx, &
                    mhd_config%ymin + mhd_config%ny * mhd_config%dy, &
                    merge(mhd_config%zmin + mhd_config%nz * mhd_config%dz, mhd_config%zmax, &
                          ndim_field == 3)]
    endif
    call cpu_time(step1)
    do tf = tmin + 1, t_end
        write(*, "(A,I0)") " Starting step ", tf
        if (single_time_frame == 0 .and. tf <= tmax_mhd) then
            write(fname, "(A,I4.4)") trim(dir_mhd_data)//'mhd_data_', tf
            call read_field_data_parallel(fname, time_interp_flag)
            call calc_fields_gradients(time_interp_flag)
        endif
        if (tf == t_start + 1 .or. inject_new_ptl) then
            if (tf <= tmax_to_inject) then
                if (inject_large_jz) then
                    call inject_particles_at_large_jz(nptl, dt, dist_flag, particle_v0, &
                        tf - t_start, inject_same_nptl, jz_min, ncells_large_jz_norm, &
                        part_box, power_index)
                else if (inject_large_absj) then
                    call inject_particles_at_large_absj(nptl, dt, dist_flag, particle_v0, &
                        tf - t_start, inject_same_nptl, absj_min, ncells_large_absj_norm, &
                        part_box, power_index)
                else if (inject_large_divv) then
                    call inject_particles_at_large_divv(nptl, dt, dist_flag, particle_v0, &
                        tf - t_start, inject_same_nptl, divv_min, ncells_large_divv_norm, &
                        part_box, power_index)
                else if (inject_large_rho) then
                    call inject_particles_at_large_rho(nptl, dt, dist_flag, particle_v0, &
                        tf - t_start, inject_same_nptl, rho_min, ncells_large_rho_norm, &
                        part_box, power_index)
                else
                    call inject_particles_spatial_uniform(nptl, dt, dist_flag, particle_v0, &
                        tf - t_start, part_box, power_index)
                endif
            endif
        endif
        if (tf == t_start + 1) then
            call distributions_diagnostics(t_start, diagnostics_directory, local_dist, &
                dump_escaped_dist)
            call quick_check(t_start, .true., diagnostics_directory)
            call get_pmax_global(t_start, .true., diagnostics_directory)
        endif
        call particle_mover(focused_transport, nlgc, kperp_kpara, nsteps_interval, &
            tf - t_start, num_fine_steps, dump_escaped_dist)
        write(*, "(A)") " Finishing moving particles "
        if (split_flag == 1) call split_particle(split_ratio, pmin_split, nsteps_interval)
        call quick_check(tf, .false., diagnostics_directory)
        call get_pmax_global(tf, .false., diagnostics_directory)
        call distributions_diagnostics(tf, diagnostics_directory, local_dist, dump_escaped_dist)
        if (time_interp_flag == 1) call copy_fields
        call cpu_time(step2)
        print '("Step ", I0, " takes ", f9.4, " seconds.")', tf, step2 - step1
        step1 = step2
    enddo
    write(*, "(A,I0,A,ES12.5,A)") " particle-steps: ", steps_total, "  mover kernel time: ", &
        kernel_ms_total * 1.0d-3, " s"
    call free_particles
    call free_tstamps_mhd
    call cpu_time(finish)
    print '("Time = ",f14.7," seconds.")', finish - start

contains

    !< Flags whose physics the GPU path does not carry stop here with a message
    subroutine check_scope
        if (restart_flag) call stop_scope("-rf (restart)")
        if (focused_transport) call stop_scope("-ft .true. (focused transport)")
        if (uniform_grid /= 1) call stop_scope("-ug 0 (non-uniform grid); pass -ug 1")
        if (spherical_coord /= 0) call stop_scope("-sc 1 (spherical coordinates)")
        if (trim(gpu_precision) == 'f32h' .and. (dpp_wave /= 0 .or. dpp_shear /= 0)) &
            call stop_scope("-dw/-ds with -gp f32h (use f32 or f64)")
        if (deltab_flag /= 0 .or. correlation_flag /= 0) call stop_scope("-db/-co maps")
        if (include_3rd_dim /= 0) call stop_scope("-i3 1")
        if (acc_by_surface /= 0) call stop_scope("-as 1")
        if (track_particle_flag) call stop_scope("-tf .true. (particle tracking)")
        if (inject_at_shock) call stop_scope("-is .true. (shock injection)")
        if (inject_large_db2) call stop_scope("-ib .true. (needs the turbulence-amplitude map)")
        if (ndim_field /= 2 .and. ndim_field /= 3) call stop_scope("-nd other than 2 or 3")
        if (size_mpi_sub /= 1) write(*, "(A)") " -sm ignored: one GPU per run"
    end subroutine check_scope

    subroutine stop_scope(what)
        character(*), intent(in) :: what
        write(*, "(A,A)") " not supported on the GPU Parker path: ", what
        stop 2
    end subroutine stop_scope

    logical function to_logical(s)
        character(*), intent(in) :: s
        select case (trim(adjustl(s)))
        case ('.true.', '.TRUE.', 'T', 't', 'true', 'True', '1')
            to_logical = .true.
        case default
            to_logical = .false.
        end select
    end function to_logical

    real(dp) function to_real(s)
        character(*), intent(in) :: s
        integer :: ios
        read(s, *, iostat=ios) to_real
        if (ios /= 0) then
            write(*, "(A,A)") " cannot read a number from ", trim(s)
            stop 1
        endif
    end function to_real

    integer function to_int(s)
        character(*), intent(in) :: s
        to_int = nint(to_real(s))
    end function to_int

    subroutine get_cmd_args
        integer :: i, n
        character(len=256) :: key, val
        n = command_argument_count()
        i = 1
        do while (i <= n)
            call get_command_argument(i, key)
            if (i + 1 > n) then
                write(*, "(A,A)") " missing value after ", trim(key)
                stop 1
            endif
            call get_command_argument(i + 1, val)
            select case (trim(key))
            case ('-qh', '--quota_hour'); quota_hour = to_real(val)
            case ('-rf', '--restart_flag'); restart_flag = to_logical(val)
            case ('-ft', '--focused_transport'); focused_transport = to_logical(val)
            case ('-nl', '--nlgc'); nlgc = to_logical(val)
            case ('-kk', '--kperp_kpara'); kperp_kpara = to_real(val)
            case ('-pv', '--particle_v0'); particle_v0 = to_real(val)
            case ('-sm', '--size_mpi_sub'); size_mpi_sub = to_int(val)
            case ('-dm', '--dir_mhd_data'); dir_mhd_data = val
            case ('-mc', '--mhd_config_filename'); mhd_config_filename = val
            case ('-nm', '--nptl_max'); nptl_max = int(to_real(val), c_int64_t)
            case ('-np', '--nptl'); nptl = to_int(val)
            case ('-ti', '--time_interp_flag'); time_interp_flag = to_int(val)
            case ('-dt', '--tinterval'); dt = to_real(val)
            case ('-ts', '--tstart'); t_start = to_int(val)
            case ('-te', '--tend'); t_end = to_int(val)
            case ('-tm', '--tmax_mhd'); tmax_mhd = to_int(val)
            case ('-st', '--single_time_frame'); single_time_frame = to_int(val)
            case ('-df', '--dist_flag'); dist_flag = to_int(val)
            case ('-pi', '--power_index'); power_index = to_real(val)
            case ('-sf', '--split_flag'); split_flag = to_int(val)
            case ('-sr', '--split_ratio'); split_ratio = to_real(val)
            case ('-ps', '--pmin_split'); pmin_split = to_real(val)
            case ('-tf', '--track_particle_flag'); track_particle_flag = to_logical(val)
            case ('-ni', '--nsteps_interval'); nsteps_interval = to_int(val)
            case ('-dd', '--diagnostics_directory'); diagnostics_directory = val
            case ('-is', '--inject_at_shock'); inject_at_shock = to_logical(val)
            case ('-in', '--inject_new_ptl'); inject_new_ptl = to_logical(val)
            case ('-ij', '--inject_large_jz'); inject_large_jz = to_logical(val)
            case ('-sn', '--inject_same_nptl'); inject_same_nptl = to_logical(val)
            case ('-jz', '--jz_min'); jz_min = to_real(val)
            case ('-nn', '--ncells_large_jz_norm'); ncells_large_jz_norm = to_int(val)
            case ('-iaj', '--inject_large_absj'); inject_large_absj = to_logical(val)
            case ('-ajm', '--absj_min'); absj_min = to_real(val)
            case ('-naj', '--ncells_large_absj_norm'); ncells_large_absj_norm = to_int(val)
            case ('-ib', '--inject_large_db2'); inject_large_db2 = to_logical(val)
            case ('-iv', '--inject_large_divv'); inject_large_divv = to_logical(val)
            case ('-dv', '--divv_min'); divv_min = to_real(val)
            case ('-nv', '--ncells_large_divv_norm'); ncells_large_divv_norm = to_int(val)
            case ('-ir', '--inject_large_rho'); inject_large_rho = to_logical(val)
            case ('-rm', '--rho_min'); rho_min = to_real(val)
            case ('-nr', '--ncells_large_rho_norm'); ncells_large_rho_norm = to_int(val)
            case ('-tti', '--tmax_to_inject'); tmax_to_inject = to_int(val)
            case ('-ip', '--inject_part_box'); inject_part_box = to_logical(val)
            case ('-xs', '--ptl_xmin'); ptl_xmin = to_real(val)
            case ('-xe', '--ptl_xmax'); ptl_xmax = to_real(val)
            case ('-ys', '--ptl_ymin'); ptl_ymin = to_real(val)
            case ('-ye', '--ptl_ymax'); ptl_ymax = to_real(val)
            case ('-zs', '--ptl_zmin'); ptl_zmin = to_real(val)
            case ('-ze', '--ptl_zmax'); ptl_zmax = to_real(val)
            case ('-cf', '--conf_file'); conf_file = val
            case ('-nf', '--num_fine_steps'); num_fine_steps = to_int(val)
            case ('-ld', '--local_dist'); local_dist = to_logical(val)
            case ('-dw', '--dpp_wave'); dpp_wave = to_int(val)
            case ('-ds', '--dpp_shear'); dpp_shear = to_int(val)
            case ('-ws', '--weak_scattering'); weak_scattering = to_int(val)
            case ('-t0', '--tau0_scattering'); tau0_scattering = to_real(val)
            case ('-db', '--deltab_flag'); deltab_flag = to_int(val)
            case ('-co', '--correlation'); correlation_flag = to_int(val)
            case ('-nd', '--ndim_field'); ndim_field = to_int(val)
            case ('-dp1', '--drift_param1'); drift_param1 = to_real(val)
            case ('-dp2', '--drift_param2'); drift_param2 = to_real(val)
            case ('-ch', '--charge'); charge = to_int(val)
            case ('-sc', '--spherical_coord'); spherical_coord = to_int(val)
            case ('-ug', '--uniform_grid'); uniform_grid = to_int(val)
            case ('-cd', '--check_drift_2d'); check_drift_2d = to_int(val)
            case ('-i3', '--include_3rd_dim'); include_3rd_dim = to_int(val)
            case ('-as', '--acc_by_surface'); acc_by_surface = to_int(val)
            case ('-vdt', '--varying_dt_mhd'); varying_dt_mhd = to_logical(val)
            case ('-ded', '--dump_escaped_dist'); dump_escaped_dist = to_logical(val)
            case ('-gp', '--gpu_precision'); gpu_precision = val
            case ('-rs', '--rng_seed'); rng_seed = int(to_real(val), c_int64_t)
            case ('-ss', '--sort_steps'); sort_steps = to_int(val)
            case ('-db2', '-nb', '-pd', '-ptf', '-de', '-du', &
                  '-sf1', '-sn1', '-s2e', '-ii', '-sf2', '-sn2')
                ! accepted and ignored: options of paths the GPU pusher does not run
            case default
                write(*, "(A,A)") " unknown option ", trim(key)
                stop 1
            end select
            i = i + 2
        enddo
        if (len_trim(dir_mhd_data) == 0) then
            write(*, "(A)") " -dm/--dir_mhd_data is required"
            stop 1
        endif
    end subroutine get_cmd_args
end program stochastic_hip
