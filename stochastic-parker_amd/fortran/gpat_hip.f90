!*******************************************************************************
!< Host side of the Parker pusher in Fortran: the module procedures the
!< reference driver calls (stochastic-mhd.f90:189-566), with the reference's
!< names and argument lists, backed by libsp_hip.so through sp_hip_binding.
!<
!< Module state mirrors the reference's module globals (mhd_config, tstamps,
!< fconfig, nptl_current, leak, leak_negp, ...); the particles and fields live
!< on the GPU inside one context.  Procedures outside the Parker pusher path
!< (focused transport, NLGC, threshold/shock injection, spherical and
!< non-uniform grids, tracking) stop with a message.
!*******************************************************************************
module gpat_hip
    use iso_c_binding
    use sp_hip_binding
    implicit none
    private
    save
    public mhd_configuration, mhd_config, tstamps_mhd, get_variable, load_mhd_config, &
        init_tstamps_mhd, calc_tstamps_mhd, load_tstamps_mhd, free_tstamps_mhd, &
        read_particle_boundary_conditions, read_particle_params, read_diagnostics_params, &
        set_drift_parameters, set_flag_check_drift_2d, set_precision, set_sort_steps, &
        set_dpp_params, set_nlgc, &
        init_particles, free_particles, init_field_data, free_field_data, &
        read_field_data_parallel, calc_fields_gradients, copy_fields, &
        inject_particles_spatial_uniform, particle_mover, split_particle, &
        inject_particles_at_large_jz, inject_particles_at_large_absj, &
        inject_particles_at_large_divv, inject_particles_at_large_rho, &
        init_particle_distributions, distributions_diagnostics, quick_check, &
        get_pmax_global, nptl_current, leak, leak_negp, steps_total, kernel_ms_total, &
        fglobal, pbins_edges_global, check_status

    integer, parameter :: dp = kind(1.0d0)

    !< mhd_config.dat record (mhd_config.f90:16-29): 13 reals, 13 integers
    type mhd_configuration
        real(dp) :: dx = 1.0_dp, dy = 1.0_dp, dz = 1.0_dp
        real(dp) :: xmin = 0.0_dp, ymin = 0.0_dp, zmin = 0.0_dp
        real(dp) :: xmax = 1.0_dp, ymax = 1.0_dp, zmax = 1.0_dp
        real(dp) :: lx = 1.0_dp, ly = 1.0_dp, lz = 1.0_dp
        real(dp) :: dt_out = 0.25_dp
        integer :: nx = 1, ny = 1, nz = 1
        integer :: nxs = 1, nys = 1, nzs = 1
        integer :: topox = 1, topoy = 1, topoz = 1
        integer :: nvar = 9
        integer :: bcx = 0, bcy = 0, bcz = 0
    end type mhd_configuration

    type(mhd_configuration) :: mhd_config
    real(dp), allocatable :: tstamps_mhd(:)
    integer :: tstart_mhd, tend_mhd

    type(sp_config), target :: cfg
    type(sp_dist_config) :: dcfg
    type(c_ptr) :: ctx = c_null_ptr
    integer :: ndim_field = 2
    integer(c_int64_t) :: nptl_current = 0
    real(dp) :: leak = 0.0_dp, leak_negp = 0.0_dp
    integer(c_int64_t) :: steps_total = 0
    real(dp) :: kernel_ms_total = 0.0_dp
    real(dp) :: p0_, pmin_, pmax_
    real(dp), allocatable :: fglobal(:), pbins_edges_global(:)
    real(dp), allocatable, target :: flocal(:, :)
    integer(c_int64_t) :: loc_size(4) = 0
    real(dp), allocatable, target :: fescaped(:), fescaped_local(:, :)
    integer(c_int64_t) :: esc_size(4) = 0
    !< Fortran-order dimensions (nmu, npbins, nrx, nry, nrz) of local set k
    integer(c_int64_t) :: loc_dims(5, 4) = 0
    real(dp), allocatable :: mubins_edges_global(:)

    !< HDF5 writer of the distribution files (fortran/sp_h5.c, libsp_h5.so)
    interface
        integer(c_int64_t) function sp_h5_open(fname, create) bind(C, name="sp_h5_open")
            import :: c_int64_t, c_int32_t, c_char
            character(kind=c_char), intent(in) :: fname(*)
            integer(c_int32_t), value :: create
        end function sp_h5_open
        integer(c_int32_t) function sp_h5_write(fid, name, rank, fdims, data) &
                bind(C, name="sp_h5_write")
            import :: c_int64_t, c_int32_t, c_char, c_double
            integer(c_int64_t), value :: fid
            character(kind=c_char), intent(in) :: name(*)
            integer(c_int32_t), value :: rank
            integer(c_int64_t), intent(in) :: fdims(*)
            real(c_double), intent(in) :: data(*)
        end function sp_h5_write
        integer(c_int32_t) function sp_h5_close(fid) bind(C, name="sp_h5_close")
            import :: c_int64_t, c_int32_t
            integer(c_int64_t), value :: fid
        end function sp_h5_close
    end interface

contains

    subroutine check_status(rc, what)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: what
        if (rc /= SP_OK) then
            write(*, "(A,A,A,I0,A,A)") " libsp_hip: ", what, " failed (", rc, "): ", &
                sp_error_message()
            stop 1
        endif
    end subroutine check_status

    subroutine unsupported(what)
        character(*), intent(in) :: what
        write(*, "(A,A)") " not supported on the GPU Parker path: ", what
        stop 2
    end subroutine unsupported

    !---------------------------------------------------------------------------
    !< conf.dat reader (read_config.f90:22-45): value after the delimiter on the
    !< first non-comment line containing var_name; -1.0 when absent.  The
    !< search restarts from the top for every key, so key order is free.
    !---------------------------------------------------------------------------
    function get_variable(fname, var_name, delimiter) result(var_value)
        character(*), intent(in) :: fname, var_name, delimiter
        real(dp) :: var_value
        character(len=256) :: line
        integer :: fh, ios, k, e
        var_value = -1.0_dp
        open(newunit=fh, file=fname, status='old', action='read', iostat=ios)
        if (ios /= 0) then
            write(*, "(A,A)") " cannot open configuration file ", trim(fname)
            stop 1
        endif
        do
            read(fh, '(A)', iostat=ios) line
            if (ios /= 0) exit
            line = adjustl(line)
            if (len_trim(line) == 0) cycle
            if (line(1:1) == '!') cycle
            if (index(line, var_name) == 0) cycle
            k = index(line, delimiter)
            if (k == 0) cycle
            e = index(line, ";")
            if (e == 0) e = len_trim(line) + 1
            read(line(k+1:e-1), *, iostat=ios) var_value
            if (ios /= 0) var_value = -1.0_dp
            exit
        enddo
        close(fh)
    end function get_variable

    !< load_mhd_config (mhd_config.f90:139-148): 13 reals then 13 integers
    subroutine load_mhd_config(filename)
        character(*), intent(in) :: filename
        integer :: fh
        real(dp) :: r(13)
        integer :: iv(13)
        open(newunit=fh, file=filename, access='stream', status='old', form='unformatted', &
             action='read')
        read(fh, pos=1) r, iv
        close(fh)
        mhd_config%dx = r(1); mhd_config%dy = r(2); mhd_config%dz = r(3)
        mhd_config%xmin = r(4); mhd_config%ymin = r(5); mhd_config%zmin = r(6)
        mhd_config%xmax = r(7); mhd_config%ymax = r(8); mhd_config%zmax = r(9)
        mhd_config%lx = r(10); mhd_config%ly = r(11); mhd_config%lz = r(12)
        mhd_config%dt_out = r(13)
        mhd_config%nx = iv(1); mhd_config%ny = iv(2); mhd_config%nz = iv(3)
        mhd_config%nxs = iv(4); mhd_config%nys = iv(5); mhd_config%nzs = iv(6)
        mhd_config%topox = iv(7); mhd_config%topoy = iv(8); mhd_config%topoz = iv(9)
        mhd_config%nvar = iv(10)
        mhd_config%bcx = iv(11); mhd_config%bcy = iv(12); mhd_config%bcz = iv(13)
    end subroutine load_mhd_config

    subroutine init_tstamps_mhd(t_start, t_end)
        integer, intent(in) :: t_start, t_end
        tstart_mhd = t_start
        tend_mhd = t_end
        if (allocated(tstamps_mhd)) deallocate(tstamps_mhd)
        allocate(tstamps_mhd(t_end - t_start + 1))
    end subroutine init_tstamps_mhd

    !< load_tstamps_mhd (mhd_config.f90:221-255): dir_mhd_data/time_stamps.dat
    !< holds 8-byte slots, the first frame of the MHD output (read as a default
    !< integer from the first slot), the last frame, then one FP64 time per
    !< frame.  Frames t_start .. min(tmax_mhd, t_end) are read; frames past
    !< tmax_mhd repeat the last interval.
    subroutine load_tstamps_mhd(t_start, t_end, tmax_mhd, dir_mhd_data)
        integer, intent(in) :: t_start, t_end, tmax_mhd
        character(*), intent(in) :: dir_mhd_data
        character(len=256) :: filename
        integer :: fh, ts_mhd, i, nread, ios
        integer(8) :: fsize
        real(dp) :: dt_last
        filename = trim(dir_mhd_data)//"time_stamps.dat"
        open(newunit=fh, file=filename, access='stream', status='old', form='unformatted', &
             action='read', iostat=ios)
        if (ios /= 0) then
            write(*, "(A,A)") "cannot open ", trim(filename)
            stop 1
        endif
        read(fh, pos=1) ts_mhd
        nread = min(tmax_mhd, t_end) - t_start + 1
        inquire(unit=fh, size=fsize)
        if (t_start < ts_mhd .or. nread < 1 .or. &
                (int(t_start - ts_mhd + 2, 8) + nread) * 8 > fsize) then
            write(*, "(A,A,A,I0,A,I0)") trim(filename), " does not hold the frames ", &
                "from ", t_start, " to ", min(tmax_mhd, t_end)
            stop 1
        endif
        read(fh, pos=(t_start - ts_mhd + 2) * 8 + 1) tstamps_mhd(1:nread)
        close(fh)
        if (tmax_mhd < t_end) then
            if (nread < 2) then
                write(*, "(A)") "time_stamps.dat: extrapolating past tmax_mhd needs two frames"
                stop 1
            endif
            dt_last = tstamps_mhd(nread) - tstamps_mhd(nread - 1)
            do i = tmax_mhd + 1, t_end
                tstamps_mhd(i - t_start + 1) = tstamps_mhd(i - t_start) + dt_last
            enddo
        endif
    end subroutine load_tstamps_mhd

    !< calc_tstamps_mhd (mhd_config.f90:263-271)
    subroutine calc_tstamps_mhd(t_start, t_end)
        integer, intent(in) :: t_start, t_end
        integer :: i
        do i = t_start, t_end
            tstamps_mhd(i - t_start + 1) = i * mhd_config%dt_out
        enddo
    end subroutine calc_tstamps_mhd

    subroutine free_tstamps_mhd
        if (allocated(tstamps_mhd)) deallocate(tstamps_mhd)
    end subroutine free_tstamps_mhd

    !< Grid and domain of one rank holding the whole snapshot
    !< (simulation_setup.f90:171-253)
    subroutine set_grid(ndim)
        integer, intent(in) :: ndim
        ndim_field = ndim
        cfg%ndim = ndim
        cfg%nx = mhd_config%nx; cfg%ny = mhd_config%ny
        cfg%nz = merge(mhd_config%nz, 1, ndim == 3)
        cfg%dx = mhd_config%dx; cfg%dy = mhd_config%dy; cfg%dz = mhd_config%dz
        cfg%xmin = mhd_config%xmin; cfg%ymin = mhd_config%ymin; cfg%zmin = mhd_config%zmin
        cfg%xmax = mhd_config%xmin + mhd_config%nx * mhd_config%dx
        cfg%ymax = mhd_config%ymin + mhd_config%ny * mhd_config%dy
        if (ndim == 3) then
            cfg%zmax = mhd_config%zmin + mhd_config%nz * mhd_config%dz
        else
            cfg%zmax = mhd_config%zmax
        endif
        cfg%lx = mhd_config%lx; cfg%ly = mhd_config%ly; cfg%lz = mhd_config%lz
    end subroutine set_grid

    !< read_particle_boundary_conditions (simulation_setup.f90:104-150)
    subroutine read_particle_boundary_conditions(conf_file)
        character(*), intent(in) :: conf_file
        cfg%pbcx = int(get_variable(conf_file, 'pbcx', '='))
        cfg%pbcy = int(get_variable(conf_file, 'pbcy', '='))
        cfg%pbcz = int(get_variable(conf_file, 'pbcz', '='))
        if (int(get_variable(conf_file, 'mpi_sizex', '=')) > 1 .or. &
            int(get_variable(conf_file, 'mpi_sizey', '=')) > 1 .or. &
            int(get_variable(conf_file, 'mpi_sizez', '=')) > 1) then
            write(*, "(A)") " mpi_size{x,y,z} > 1 ignored: one GPU holds the whole domain"
        endif
    end subroutine read_particle_boundary_conditions

    !< read_particle_params (particle_module.f90:2778-2878)
    subroutine read_particle_params(conf_file)
        character(*), intent(in) :: conf_file
        cfg%b0 = get_variable(conf_file, 'b0', '=')
        cfg%p0 = get_variable(conf_file, 'p0', '=')
        cfg%pmin = get_variable(conf_file, 'pmin', '=')
        cfg%pmax = get_variable(conf_file, 'pmax', '=')
        cfg%momentum_dependency = int(get_variable(conf_file, 'momentum_dependency', '='))
        cfg%gamma_turb = get_variable(conf_file, 'gamma_turb', '=')
        cfg%mag_dependency = int(get_variable(conf_file, 'mag_dependency', '='))
        cfg%kpara0 = get_variable(conf_file, 'kpara0', '=')
        cfg%kret = get_variable(conf_file, 'kret', '=')
        cfg%dt_min_rel = get_variable(conf_file, 'dt_min_rel', '=')
        cfg%dt_max_rel = get_variable(conf_file, 'dt_max_rel', '=')
        cfg%acc_region_flag = int(get_variable(conf_file, 'acc_region_flag', '='))
        if (cfg%acc_region_flag == 1) then
            cfg%acc_region(1) = get_variable(conf_file, 'acc_xmin', '=')
            cfg%acc_region(2) = get_variable(conf_file, 'acc_xmax', '=')
            cfg%acc_region(3) = get_variable(conf_file, 'acc_ymin', '=')
            cfg%acc_region(4) = get_variable(conf_file, 'acc_ymax', '=')
            cfg%acc_region(5) = get_variable(conf_file, 'acc_zmin', '=')
            cfg%acc_region(6) = get_variable(conf_file, 'acc_zmax', '=')
        else
            cfg%acc_region_flag = 0
            cfg%acc_region = [0.0_dp, 1.0_dp, 0.0_dp, 1.0_dp, 0.0_dp, 1.0_dp]
        endif
        p0_ = cfg%p0
        pmin_ = cfg%pmin
        pmax_ = cfg%pmax
    end subroutine read_particle_params

    !< read_diagnostics_params (diagnostics.f90:2049-2222), Parker transport:
    !< nmu_global and nmu1..4 are 1; a set is dumped when dump_interval < nframes
    subroutine read_diagnostics_params(conf_file, nframes, local_dist)
        character(*), intent(in) :: conf_file
        integer, intent(in) :: nframes
        logical, intent(in) :: local_dist
        integer :: k
        character(len=1) :: s
        dcfg%npp_global = int(get_variable(conf_file, 'npp_global', '='))
        dcfg%nmu_global = 1
        dcfg%local_dist = merge(1, 0, local_dist)
        do k = 1, 4
            write(s, "(I1)") k
            dcfg%local(k)%enabled = merge(1, 0, &
                int(get_variable(conf_file, 'dump_interval'//s, '=')) < nframes)
            dcfg%local(k)%pmin = get_variable(conf_file, 'pmin'//s, '=')
            dcfg%local(k)%pmax = get_variable(conf_file, 'pmax'//s, '=')
            dcfg%local(k)%npbins = int(get_variable(conf_file, 'npbins'//s, '='))
            dcfg%local(k)%nmu = 1
            dcfg%local(k)%rx = int(get_variable(conf_file, 'rx'//s, '='))
            dcfg%local(k)%ry = int(get_variable(conf_file, 'ry'//s, '='))
            dcfg%local(k)%rz = int(get_variable(conf_file, 'rz'//s, '='))
            if (ndim_field < 3) dcfg%local(k)%rz = 1
        enddo
    end subroutine read_diagnostics_params

    !< set_drift_parameters (particle_module.f90:316-325)
    subroutine set_drift_parameters(drift_param1, drift_param2, charge)
        real(dp), intent(in) :: drift_param1, drift_param2
        integer, intent(in) :: charge
        cfg%drift1 = drift_param1
        cfg%drift2 = drift_param2
        cfg%charge = charge
    end subroutine set_drift_parameters

    !< set_flag_check_drift_2d (particle_module.f90:327-335)
    subroutine set_flag_check_drift_2d(check_drift_2d_flag)
        integer, intent(in) :: check_drift_2d_flag
        cfg%check_drift_2d = merge(1, 0, check_drift_2d_flag == 1)
    end subroutine set_flag_check_drift_2d

    subroutine set_precision(prec, seed)
        integer, intent(in) :: prec
        integer(c_int64_t), intent(in) :: seed
        cfg%precision = prec
        cfg%seed = seed
    end subroutine set_precision

    !< set_dpp_params (particle_module.f90:260-272): momentum diffusion by wave
    !< scattering and by flow shear in the push
    subroutine set_dpp_params(dpp_wave_int, dpp_shear_int, weak_scattering_flag, &
            tau0_scattering)
        integer, intent(in) :: dpp_wave_int, dpp_shear_int, weak_scattering_flag
        real(dp), intent(in) :: tau0_scattering
        cfg%dpp_wave = merge(1, 0, dpp_wave_int == 1)
        cfg%dpp_shear = merge(1, 0, dpp_shear_int == 1)
        cfg%weak_scattering = merge(1, 0, weak_scattering_flag == 1)
        cfg%tau0 = tau0_scattering
    end subroutine set_dpp_params

    !< The kperp model of particle_mover (particle_module.f90:1846-1851): NLGC
    !< with kperp/kpara = kperp_kpara at p0, or the standard kret ratio.  The
    !< context is created with it; particle_mover checks that its arguments agree.
    subroutine set_nlgc(nlgc, kperp_kpara)
        logical, intent(in) :: nlgc
        real(dp), intent(in) :: kperp_kpara
        cfg%nlgc = merge(1, 0, nlgc)
        cfg%kperp_kpara = kperp_kpara
    end subroutine set_nlgc

    !< spatial sort cadence inside one mover call (steps per particle; 0: once)
    subroutine set_sort_steps(nsteps)
        integer, intent(in) :: nsteps
        cfg%sort_steps = nsteps
    end subroutine set_sort_steps

    !< init_field_data (mhd_data_parallel.f90:66-101); the context is created
    !< here, after the grid and the particle parameters are known
    subroutine init_field_data(time_interp_flag, ndim)
        integer, intent(in) :: time_interp_flag, ndim
        call set_grid(ndim)
        cfg%time_interp = merge(1, 0, time_interp_flag == 1)
    end subroutine init_field_data

    !< init_particles (particle_module.f90:171-241)
    subroutine init_particles(nptl_max_allowed)
        integer(c_int64_t), intent(in) :: nptl_max_allowed
        cfg%nptl_max = nptl_max_allowed
        cfg%stream = 0
        cfg%sort_every = 1
        cfg%reserved0 = 0
        call check_status(sp_ctx_create(cfg, ctx), "sp_ctx_create")
        nptl_current = 0
        leak = 0.0_dp
        leak_negp = 0.0_dp
    end subroutine init_particles

    subroutine free_particles
        integer(c_int) :: rc
        if (c_associated(ctx)) rc = sp_ctx_destroy(ctx)
        ctx = c_null_ptr
    end subroutine free_particles

    subroutine free_field_data
    end subroutine free_field_data

    !< read_field_data_parallel (mhd_data_parallel.f90:224-297): the whole
    !< ghosted snapshot (8, nxg, nyg, nzg) in FP32, read and handed to the GPU
    subroutine read_field_data_parallel(filename, var_flag)
        character(*), intent(in) :: filename
        integer, intent(in) :: var_flag
        real(c_float), allocatable :: buf(:)
        integer(c_int64_t) :: nvals
        integer :: fh, nxg, nyg, nzg
        nxg = cfg%nx + 4
        nyg = cfg%ny + 4
        nzg = merge(cfg%nz + 4, cfg%nz, ndim_field == 3)
        nvals = 8_c_int64_t * nxg * nyg * nzg
        allocate(buf(nvals))
        open(newunit=fh, file=filename, access='stream', status='old', form='unformatted', &
             action='read')
        read(fh, pos=1) buf
        close(fh)
        call check_status(sp_fields_upload(ctx, int(var_flag, c_int32_t), buf, nvals), &
            "sp_fields_upload")
        deallocate(buf)
    end subroutine read_field_data_parallel

    !< calc_fields_gradients (mhd_data_parallel.f90:504-605), on the GPU
    subroutine calc_fields_gradients(var_flag)
        integer, intent(in) :: var_flag
        call check_status(sp_fields_gradients(ctx, int(var_flag, c_int32_t), c_null_ptr), &
            "sp_fields_gradients")
    end subroutine calc_fields_gradients

    !< copy_fields (mhd_data_parallel.f90:1920-1923)
    subroutine copy_fields
        call check_status(sp_fields_copy(ctx), "sp_fields_copy")
    end subroutine copy_fields

    !< inject_particles_spatial_uniform (particle_module.f90:454-530)
    subroutine inject_particles_spatial_uniform(nptl, dt, dist_flag, particle_v0, ct_mhd, &
            part_box, power_index)
        integer, intent(in) :: nptl, dist_flag, ct_mhd
        real(dp), intent(in) :: dt, particle_v0, power_index
        real(dp), intent(in) :: part_box(6)
        real(dp) :: t0, dtf
        integer(c_int32_t) :: tag_max
        integer(c_int64_t) :: ns, ne
        t0 = tstamps_mhd(ct_mhd)
        dtf = tstamps_mhd(ct_mhd + 1) - t0
        call check_status(sp_inject_particles_spatial_uniform(ctx, int(nptl, c_int64_t), dt, &
            int(dist_flag, c_int32_t), particle_v0, t0, dtf, part_box, power_index), &
            "sp_inject_particles_spatial_uniform")
        call check_status(sp_particles_counts(ctx, nptl_current, ns, ne, tag_max), &
            "sp_particles_counts")
    end subroutine inject_particles_spatial_uniform

    !< inject_particles_at_large_jz (particle_module.f90:785-903)
    subroutine inject_particles_at_large_jz(nptl, dt, dist_flag, particle_v0, ct_mhd, &
            inject_same_nptl, jz_min, ncells_large_jz_norm, part_box, power_index)
        integer, intent(in) :: nptl, dist_flag, ct_mhd, ncells_large_jz_norm
        logical, intent(in) :: inject_same_nptl
        real(dp), intent(in) :: dt, jz_min, power_index, particle_v0
        real(dp), intent(in) :: part_box(6)
        call inject_threshold(SP_INJECT_JZ, "jz", nptl, dt, dist_flag, particle_v0, ct_mhd, &
            inject_same_nptl, jz_min, ncells_large_jz_norm, part_box, power_index)
    end subroutine inject_particles_at_large_jz

    !< inject_particles_at_large_absj (particle_module.f90:922-1056)
    subroutine inject_particles_at_large_absj(nptl, dt, dist_flag, particle_v0, ct_mhd, &
            inject_same_nptl, absj_min, ncells_large_absj_norm, part_box, power_index)
        integer, intent(in) :: nptl, dist_flag, ct_mhd, ncells_large_absj_norm
        logical, intent(in) :: inject_same_nptl
        real(dp), intent(in) :: dt, absj_min, power_index, particle_v0
        real(dp), intent(in) :: part_box(6)
        call inject_threshold(SP_INJECT_ABSJ, "absj", nptl, dt, dist_flag, particle_v0, ct_mhd, &
            inject_same_nptl, absj_min, ncells_large_absj_norm, part_box, power_index)
    end subroutine inject_particles_at_large_absj

    !< inject_particles_at_large_divv (particle_module.f90:1205-1337)
    subroutine inject_particles_at_large_divv(nptl, dt, dist_flag, particle_v0, ct_mhd, &
            inject_same_nptl, divv_min, ncells_large_divv_norm, part_box, power_index)
        integer, intent(in) :: nptl, dist_flag, ct_mhd, ncells_large_divv_norm
        logical, intent(in) :: inject_same_nptl
        real(dp), intent(in) :: dt, divv_min, power_index, particle_v0
        real(dp), intent(in) :: part_box(6)
        call inject_threshold(SP_INJECT_DIVV, "divv", nptl, dt, dist_flag, particle_v0, ct_mhd, &
            inject_same_nptl, divv_min, ncells_large_divv_norm, part_box, power_index)
    end subroutine inject_particles_at_large_divv

    !< inject_particles_at_large_rho (particle_module.f90:1356-1468)
    subroutine inject_particles_at_large_rho(nptl, dt, dist_flag, particle_v0, ct_mhd, &
            inject_same_nptl, rho_min, ncells_large_rho_norm, part_box, power_index)
        integer, intent(in) :: nptl, dist_flag, ct_mhd, ncells_large_rho_norm
        logical, intent(in) :: inject_same_nptl
        real(dp), intent(in) :: dt, rho_min, power_index, particle_v0
        real(dp), intent(in) :: part_box(6)
        call inject_threshold(SP_INJECT_RHO, "rho", nptl, dt, dist_flag, particle_v0, ct_mhd, &
            inject_same_nptl, rho_min, ncells_large_rho_norm, part_box, power_index)
    end subroutine inject_particles_at_large_rho

    subroutine inject_threshold(quantity, what, nptl, dt, dist_flag, particle_v0, ct_mhd, &
            inject_same_nptl, threshold, ncells_norm, part_box, power_index)
        integer(c_int32_t), intent(in) :: quantity
        character(*), intent(in) :: what
        integer, intent(in) :: nptl, dist_flag, ct_mhd, ncells_norm
        logical, intent(in) :: inject_same_nptl
        real(dp), intent(in) :: dt, threshold, power_index, particle_v0
        real(dp), intent(in) :: part_box(6)
        real(dp) :: t0, dtf
        integer(c_int32_t) :: tag_max
        integer(c_int64_t) :: ns, ne, ninj
        t0 = tstamps_mhd(ct_mhd)
        dtf = tstamps_mhd(ct_mhd + 1) - t0
        call check_status(sp_inject_particles_threshold(ctx, quantity, int(nptl, c_int64_t), dt, &
            int(dist_flag, c_int32_t), particle_v0, t0, dtf, merge(1, 0, inject_same_nptl), &
            threshold, int(ncells_norm, c_int64_t), part_box, power_index, ninj), &
            "sp_inject_particles_threshold")
        call check_status(sp_particles_counts(ctx, nptl_current, ns, ne, tag_max), &
            "sp_particles_counts")
        write(*, "(A,A,A,I0,A)") "Finished injecting particles where ", what, " is large (", &
            ninj, ")"
    end subroutine inject_threshold

    !< particle_mover (particle_module.f90:1846-1974)
    subroutine particle_mover(focused_transport, nlgc, kperp_kpara, nsteps_interval, &
            mhd_tframe, num_fine_steps, dump_escaped_dist)
        logical, intent(in) :: focused_transport, nlgc, dump_escaped_dist
        real(dp), intent(in) :: kperp_kpara
        integer, intent(in) :: nsteps_interval, mhd_tframe, num_fine_steps
        type(sp_mover_stats) :: st
        real(dp) :: t0, dtf
        if (focused_transport) call unsupported("focused transport (-ft .true.)")
        if (merge(1, 0, nlgc) /= cfg%nlgc .or. (nlgc .and. kperp_kpara /= cfg%kperp_kpara)) &
            call unsupported("changing -nl / -kk after init_particles (set_nlgc)")
        t0 = tstamps_mhd(mhd_tframe)
        dtf = tstamps_mhd(mhd_tframe + 1) - t0
        call check_status(sp_particle_mover(ctx, t0, dtf, int(num_fine_steps, c_int32_t), &
            int(nsteps_interval, c_int32_t), st), "sp_particle_mover")
        nptl_current = st%nptl_current
        leak = st%leak
        leak_negp = st%leak_negp
        steps_total = steps_total + st%steps
        kernel_ms_total = kernel_ms_total + st%kernel_ms
    end subroutine particle_mover

    !< split_particle (particle_module.f90:5430-5480)
    subroutine split_particle(split_ratio, pmin_split, nsteps_interval)
        real(dp), intent(in) :: split_ratio, pmin_split
        integer, intent(in) :: nsteps_interval
        integer(c_int32_t) :: tag_max
        integer(c_int64_t) :: ns, ne
        call check_status(sp_split_particle(ctx, split_ratio, pmin_split), "sp_split_particle")
        call check_status(sp_particles_counts(ctx, nptl_current, ns, ne, tag_max), &
            "sp_particles_counts")
    end subroutine split_particle

    !< init_particle_distributions (diagnostics.f90:178-226)
    subroutine init_particle_distributions(local_dist, dump_escaped_dist)
        logical, intent(in) :: local_dist, dump_escaped_dist
        integer :: i, k
        integer(c_int64_t) :: dims(5)
        real(dp) :: pmin_log, dp_log
        dcfg%escaped_dist = merge(1, 0, dump_escaped_dist)
        call check_status(sp_distributions_init(ctx, dcfg), "sp_distributions_init")
        allocate(fglobal(dcfg%npp_global), pbins_edges_global(dcfg%npp_global + 1))
        allocate(mubins_edges_global(dcfg%nmu_global + 1))
        pmin_log = log10(pmin_)
        dp_log = (log10(pmax_) - pmin_log) / dcfg%npp_global
        do i = 1, dcfg%npp_global + 1
            pbins_edges_global(i) = 10**(pmin_log + (i-1) * dp_log)
        enddo
        call mu_edges(dcfg%nmu_global, mubins_edges_global)
        do k = 1, 4
            call check_status(sp_local_dist_dims(ctx, int(k, c_int32_t), dims), &
                "sp_local_dist_dims")
            loc_dims(:, k) = dims
            loc_size(k) = product(dims)
            if (dims(1) == 0) loc_size(k) = 0
            ! init_local_escaped_distributions (diagnostics.f90:359-395): the
            ! x, y[, z] face arrays of set k, (nmu, npbins, n_b, n_c, 2) each
            esc_size(k) = 0
            if (dump_escaped_dist .and. loc_size(k) > 0) then
                esc_size(k) = dims(1) * dims(2) * 2 * (dims(4) * dims(5) + dims(3) * dims(5))
                if (ndim_field > 2) esc_size(k) = esc_size(k) + dims(1) * dims(2) * 2 * dims(3) * dims(4)
            endif
        enddo
        allocate(flocal(max(1_c_int64_t, maxval(loc_size)), 4))
        if (dump_escaped_dist) then
            allocate(fescaped(dcfg%nmu_global * dcfg%npp_global * 2 * ndim_field))
            allocate(fescaped_local(max(1_c_int64_t, maxval(esc_size)), 4))
        endif
    end subroutine init_particle_distributions

    !< mubins edges -1 + (i - 1) dmu with dmu = 2.0 / nmu in default real
    !< (diagnostics.f90:206-209)
    subroutine mu_edges(nmu, edges)
        integer, intent(in) :: nmu
        real(dp), intent(out) :: edges(:)
        real(dp) :: dmu
        integer :: i
        dmu = 2.0 / nmu
        do i = 1, nmu + 1
            edges(i) = -1.0 + (i - 1) * dmu
        enddo
    end subroutine mu_edges

    !< Opens (create) or reopens the HDF5 file fname; stops on failure.
    function h5_file(fname, create) result(fid)
        character(*), intent(in) :: fname
        logical, intent(in) :: create
        integer(c_int64_t) :: fid
        fid = sp_h5_open(trim(fname)//c_null_char, merge(1_c_int32_t, 0_c_int32_t, create))
        if (fid < 0) then
            write(*, "(A,A)") "HDF5: cannot create ", trim(fname)
            stop 1
        endif
    end function h5_file

    !< One FP64 dataset with Fortran-order dimensions dims.
    subroutine h5_dataset(fid, name, dims, data)
        integer(c_int64_t), intent(in) :: fid
        character(*), intent(in) :: name
        integer(c_int64_t), intent(in) :: dims(:)
        real(dp), intent(in) :: data(*)
        if (sp_h5_write(fid, trim(name)//c_null_char, int(size(dims), c_int32_t), dims, &
                data) /= 0) then
            write(*, "(A,A)") "HDF5: cannot write dataset ", trim(name)
            stop 1
        endif
    end subroutine h5_dataset

    subroutine h5_done(fid)
        integer(c_int64_t), intent(in) :: fid
        if (sp_h5_close(fid) /= 0) then
            write(*, "(A)") "HDF5: closing a file failed"
            stop 1
        endif
    end subroutine h5_done

    !< save_mu_pbins (diagnostics.f90:1244-1278): the mu and momentum bin edges
    subroutine save_mu_pbins(fid, nmu, npp, dname_mubins, dname_pbins, pmin, pmax)
        integer(c_int64_t), intent(in) :: fid
        integer, intent(in) :: nmu, npp
        character(*), intent(in) :: dname_mubins, dname_pbins
        real(dp), intent(in) :: pmin, pmax
        real(dp) :: mue(nmu + 1), pbe(npp + 1), pmin_log, dp_log
        integer :: i
        call mu_edges(nmu, mue)
        pmin_log = log10(pmin)
        dp_log = (log10(pmax) - pmin_log) / npp
        do i = 1, npp + 1
            pbe(i) = 10**(pmin_log + (i-1) * dp_log)
        enddo
        call h5_dataset(fid, dname_mubins, [int(nmu + 1, c_int64_t)], mue)
        call h5_dataset(fid, dname_pbins, [int(npp + 1, c_int64_t)], pbe)
    end subroutine save_mu_pbins

    !< distributions_diagnostics (diagnostics.f90:1652-1682): f(p) and the local
    !< distributions into fdists_NNNN.h5 (save_global_distributions,
    !< save_local_distributions, :1285-1323, :1420-1488), the distributions of
    !< the particles that escaped since the previous frame into
    !< escaped_dists_NNNN.h5 (save_escaped_distributions,
    !< save_local_escaped_distributions, :1331-1372, :1533-1640): the
    !< reference's dataset names, Fortran-order shapes and bin edges
    subroutine distributions_diagnostics(iframe, file_path, local_dist, dump_escaped_dist)
        integer, intent(in) :: iframe
        character(*), intent(in) :: file_path
        logical, intent(in) :: local_dist, dump_escaped_dist
        type(c_ptr) :: lp(4)
        character(len=256) :: fname
        character(len=1) :: s
        integer(c_int64_t) :: fid, d(5), off
        integer :: k
        do k = 1, 4
            lp(k) = c_null_ptr
            if (local_dist .and. loc_size(k) > 0) lp(k) = c_loc(flocal(1, k))
        enddo
        call check_status(sp_distributions(ctx, c_loc_dp(fglobal), lp(1), lp(2), lp(3), lp(4)), &
            "sp_distributions")
        write(fname, "(A,A,I4.4,A)") trim(file_path), "fdists_", iframe, ".h5"
        fid = h5_file(fname, .true.)
        call save_mu_pbins(fid, dcfg%nmu_global, dcfg%npp_global, "mubins_edges_global", &
            "pbins_edges_global", pmin_, pmax_)
        call h5_dataset(fid, "fglobal", [int(dcfg%nmu_global, c_int64_t), &
            int(dcfg%npp_global, c_int64_t)], fglobal)
        do k = 1, 4
            if (local_dist .and. loc_size(k) > 0) then
                write(s, "(I1)") k
                call save_mu_pbins(fid, int(loc_dims(1, k)), int(loc_dims(2, k)), &
                    "mubins"//s//"_edges", "pbins"//s//"_edges", dcfg%local(k)%pmin, &
                    dcfg%local(k)%pmax)
                call h5_dataset(fid, "flocal"//s, loc_dims(:, k), flocal(1:loc_size(k), k))
            endif
        enddo
        call h5_done(fid)
        if (.not. dump_escaped_dist) return
        ! calc_escaped_distributions (diagnostics.f90:913-1231) happened in the
        ! mover's removal passes; the face arrays of set k come back to back
        do k = 1, 4
            lp(k) = c_null_ptr
            if (esc_size(k) > 0) lp(k) = c_loc(fescaped_local(1, k))
        enddo
        call check_status(sp_escaped_distributions(ctx, c_loc_dp(fescaped), lp(1), lp(2), &
            lp(3), lp(4)), "sp_escaped_distributions")
        write(fname, "(A,A,I4.4,A)") trim(file_path), "escaped_dists_", iframe, ".h5"
        fid = h5_file(fname, .true.)
        call save_mu_pbins(fid, dcfg%nmu_global, dcfg%npp_global, "mubins_edges", &
            "pbins_edges", pmin_, pmax_)
        call h5_dataset(fid, "fescaped", [int(dcfg%nmu_global, c_int64_t), &
            int(dcfg%npp_global, c_int64_t), int(2 * ndim_field, c_int64_t)], fescaped)
        do k = 1, 4
            if (esc_size(k) == 0) cycle
            write(s, "(I1)") k
            d = loc_dims(:, k)
            call save_mu_pbins(fid, int(d(1)), int(d(2)), "mubins"//s//"_edges", &
                "pbins"//s//"_edges", dcfg%local(k)%pmin, dcfg%local(k)%pmax)
            call h5_dataset(fid, "fescaped"//s//"_x", [d(1), d(2), d(4), d(5), 2_c_int64_t], &
                fescaped_local(1:, k))
            off = d(1) * d(2) * d(4) * d(5) * 2
            call h5_dataset(fid, "fescaped"//s//"_y", [d(1), d(2), d(3), d(5), 2_c_int64_t], &
                fescaped_local(off + 1:, k))
            if (ndim_field > 2) then
                off = off + d(1) * d(2) * d(3) * d(5) * 2
                call h5_dataset(fid, "fescaped"//s//"_z", [d(1), d(2), d(3), d(4), 2_c_int64_t], &
                    fescaped_local(off + 1:, k))
            endif
        enddo
        call h5_done(fid)
    end subroutine distributions_diagnostics

    function c_loc_dp(a) result(p)
        real(dp), target, intent(in) :: a(:)
        type(c_ptr) :: p
        p = c_loc(a(1))
    end function c_loc_dp

    !< quick_check (diagnostics.f90:116-170), same quick.dat columns and format
    subroutine quick_check(iframe, if_create_file, file_path)
        integer, intent(in) :: iframe
        logical, intent(in) :: if_create_file
        character(*), intent(in) :: file_path
        real(dp) :: q(9)
        integer :: fh
        call check_status(sp_quick_check(ctx, q), "sp_quick_check")
        if (if_create_file) then
            open(newunit=fh, file=trim(file_path)//'quick.dat', status='replace')
            write(fh, "(A6,8A13)") "iframe", "nptl_current", "nptl_split", &
                "ntot", "leak", "leak_negp", "pdt_min", "pdt_max", "pdt_avg"
        else
            open(newunit=fh, file=trim(file_path)//'quick.dat', status="old", &
                position="append", action="write")
        endif
        write(fh, "(I6.6,8E13.6)") iframe, q(1:8)
        close(fh)
    end subroutine quick_check

    !< get_pmax_global (diagnostics.f90:1691-1719)
    subroutine get_pmax_global(iframe, if_create_file, file_path)
        integer, intent(in) :: iframe
        logical, intent(in) :: if_create_file
        character(*), intent(in) :: file_path
        real(dp) :: q(9)
        integer :: fh
        call check_status(sp_quick_check(ctx, q), "sp_quick_check")
        if (if_create_file) then
            open(newunit=fh, file=trim(file_path)//'pmax_global.dat', status='replace')
        else
            open(newunit=fh, file=trim(file_path)//'pmax_global.dat', status="old", &
                position="append", action="write")
        endif
        write(fh, "(E13.6)") q(9)
        close(fh)
    end subroutine get_pmax_global
end module gpat_hip
