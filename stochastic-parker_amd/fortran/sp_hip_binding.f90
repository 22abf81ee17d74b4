!*******************************************************************************
!< iso_c_binding interfaces of libsp_hip.so (include/sp_hip.h).
!< Derived types mirror the C structs field by field (bind(C) gives the C
!< layout); the unsigned C members (seed, stream) are carried as signed
!< integers of the same width.
!*******************************************************************************
module sp_hip_binding
    use iso_c_binding
    implicit none
    private
    public sp_config, sp_particles, sp_mover_stats, sp_local_dist, sp_dist_config, sp_variant, &
        sp_tune_result
    public sp_ctx_create, sp_ctx_destroy, sp_last_error, sp_abi_version, &
        sp_fields_upload, sp_fields_gradients, sp_fields_copy, &
        sp_particles_upload, sp_particles_download, sp_particles_counts, &
        sp_inject_particles_spatial_uniform, sp_ncells_above_threshold, &
        sp_inject_particles_threshold, sp_particle_mover, sp_sort_particles, &
        sp_split_particle, sp_step_once, sp_distributions_init, sp_distributions, &
        sp_local_dist_dims, sp_escaped_distributions, sp_quick_check, sp_calibration_copy, &
        sp_error_message, &
        sp_set_variants, sp_inject_variants, sp_variant_distributions, sp_autotune_mover
    public SP_OK, SP_PREC_F64, SP_PREC_F32, SP_PREC_F32H
    public SP_INJECT_JZ, SP_INJECT_ABSJ, SP_INJECT_DIVV, SP_INJECT_RHO

    integer(c_int), parameter :: SP_OK = 0
    integer(c_int32_t), parameter :: SP_PREC_F64 = 0
    integer(c_int32_t), parameter :: SP_PREC_F32 = 1
    integer(c_int32_t), parameter :: SP_PREC_F32H = 2
    integer(c_int32_t), parameter :: SP_INJECT_JZ = 1, SP_INJECT_ABSJ = 2
    integer(c_int32_t), parameter :: SP_INJECT_DIVV = 4, SP_INJECT_RHO = 5

    type, bind(C) :: sp_config
        integer(c_int32_t) :: ndim, nx, ny, nz
        real(c_double) :: dx, dy, dz
        real(c_double) :: xmin, ymin, zmin
        real(c_double) :: xmax, ymax, zmax
        real(c_double) :: lx, ly, lz
        integer(c_int32_t) :: pbcx, pbcy, pbcz
        integer(c_int32_t) :: momentum_dependency, mag_dependency, charge
        real(c_double) :: p0, pmin, pmax, b0
        real(c_double) :: gamma_turb
        real(c_double) :: kpara0, kret
        real(c_double) :: dt_min_rel, dt_max_rel
        real(c_double) :: drift1, drift2
        integer(c_int32_t) :: check_drift_2d, acc_region_flag
        real(c_double) :: acc_region(6)
        integer(c_int32_t) :: time_interp, precision
        integer(c_int64_t) :: seed
        integer(c_int32_t) :: stream, sort_every, sort_steps, reserved0
        integer(c_int64_t) :: nptl_max
        integer(c_int32_t) :: dpp_wave, dpp_shear, weak_scattering, nlgc
        real(c_double) :: tau0, kperp_kpara
    end type sp_config

    type, bind(C) :: sp_particles
        type(c_ptr) :: x, y, z, p, v, mu, weight, t, dt
        type(c_ptr) :: split_times, count_flag
        type(c_ptr) :: origin, nsteps_tracked, nsteps_pushed, tag_injected, tag_splitted
        type(c_ptr) :: rng_step
    end type sp_particles

    type, bind(C) :: sp_mover_stats
        integer(c_int64_t) :: steps, nptl_current, nptl_escaped
        real(c_double) :: leak, leak_negp, kernel_ms, total_ms
    end type sp_mover_stats

    type, bind(C) :: sp_local_dist
        integer(c_int32_t) :: enabled, npbins, nmu, rx, ry, rz
        real(c_double) :: pmin, pmax
    end type sp_local_dist

    type, bind(C) :: sp_variant
        real(c_double) :: kpara0, kret, gamma_turb, reserved0
    end type sp_variant

    type, bind(C) :: sp_tune_result
        integer(c_int32_t) :: blocks_per_cu, sort_steps, ntrials, particles_per_lane
        real(c_double) :: steps_per_s, default_steps_per_s
    end type sp_tune_result

    type, bind(C) :: sp_dist_config
        integer(c_int32_t) :: npp_global, nmu_global, local_dist, escaped_dist
        type(sp_local_dist) :: local(4)
    end type sp_dist_config

    interface
        integer(c_int) function sp_ctx_create(cfg, ctx) bind(C, name="sp_ctx_create")
            import :: c_int, c_ptr, sp_config
            type(sp_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: ctx
        end function sp_ctx_create

        integer(c_int) function sp_ctx_destroy(ctx) bind(C, name="sp_ctx_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function sp_ctx_destroy

        type(c_ptr) function sp_last_error() bind(C, name="sp_last_error")
            import :: c_ptr
        end function sp_last_error

        integer(c_int) function sp_abi_version() bind(C, name="sp_abi_version")
            import :: c_int
        end function sp_abi_version

        integer(c_int) function sp_fields_upload(ctx, slot, nodes8, nvals) &
                bind(C, name="sp_fields_upload")
            import :: c_int, c_int32_t, c_int64_t, c_ptr, c_float
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: slot
            real(c_float), intent(in) :: nodes8(*)
            integer(c_int64_t), value :: nvals
        end function sp_fields_upload

        integer(c_int) function sp_fields_gradients(ctx, slot, farray32) &
                bind(C, name="sp_fields_gradients")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: slot
            type(c_ptr), value :: farray32
        end function sp_fields_gradients

        integer(c_int) function sp_fields_copy(ctx) bind(C, name="sp_fields_copy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function sp_fields_copy

        integer(c_int) function sp_particles_upload(ctx, host, n) &
                bind(C, name="sp_particles_upload")
            import :: c_int, c_int64_t, c_ptr, sp_particles
            type(c_ptr), value :: ctx
            type(sp_particles), intent(in) :: host
            integer(c_int64_t), value :: n
        end function sp_particles_upload

        integer(c_int) function sp_particles_download(ctx, host, capacity, n) &
                bind(C, name="sp_particles_download")
            import :: c_int, c_int64_t, c_ptr, sp_particles
            type(c_ptr), value :: ctx
            type(sp_particles), intent(inout) :: host
            integer(c_int64_t), value :: capacity
            integer(c_int64_t), intent(out) :: n
        end function sp_particles_download

        integer(c_int) function sp_particles_counts(ctx, nptl_current, nptl_split, &
                nptl_escaped, tag_max) bind(C, name="sp_particles_counts")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: nptl_current, nptl_split, nptl_escaped
            integer(c_int32_t), intent(out) :: tag_max
        end function sp_particles_counts

        integer(c_int) function sp_inject_particles_spatial_uniform(ctx, nptl, dt, &
                dist_flag, particle_v0, t0, dtf, part_box, power_index) &
                bind(C, name="sp_inject_particles_spatial_uniform")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: nptl
            real(c_double), value :: dt
            integer(c_int32_t), value :: dist_flag
            real(c_double), value :: particle_v0, t0, dtf
            real(c_double), intent(in) :: part_box(6)
            real(c_double), value :: power_index
        end function sp_inject_particles_spatial_uniform

        integer(c_int) function sp_ncells_above_threshold(ctx, quantity, threshold, part_box, &
                ncells) bind(C, name="sp_ncells_above_threshold")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: quantity
            real(c_double), value :: threshold
            real(c_double), intent(in) :: part_box(6)
            integer(c_int64_t), intent(out) :: ncells
        end function sp_ncells_above_threshold

        integer(c_int) function sp_inject_particles_threshold(ctx, quantity, nptl, dt, &
                dist_flag, particle_v0, t0, dtf, inject_same_nptl, threshold, ncells_norm, &
                part_box, power_index, nptl_injected) &
                bind(C, name="sp_inject_particles_threshold")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: quantity
            integer(c_int64_t), value :: nptl
            real(c_double), value :: dt
            integer(c_int32_t), value :: dist_flag
            real(c_double), value :: particle_v0, t0, dtf
            integer(c_int32_t), value :: inject_same_nptl
            real(c_double), value :: threshold
            integer(c_int64_t), value :: ncells_norm
            real(c_double), intent(in) :: part_box(6)
            real(c_double), value :: power_index
            integer(c_int64_t), intent(out) :: nptl_injected
        end function sp_inject_particles_threshold

        integer(c_int) function sp_particle_mover(ctx, t0, dtf, num_fine_steps, &
                nsteps_interval, stats) bind(C, name="sp_particle_mover")
            import :: c_int, c_int32_t, c_double, c_ptr, sp_mover_stats
            type(c_ptr), value :: ctx
            real(c_double), value :: t0, dtf
            integer(c_int32_t), value :: num_fine_steps, nsteps_interval
            type(sp_mover_stats), intent(out) :: stats
        end function sp_particle_mover

        integer(c_int) function sp_sort_particles(ctx) bind(C, name="sp_sort_particles")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function sp_sort_particles

        integer(c_int) function sp_split_particle(ctx, split_ratio, pmin_split) &
                bind(C, name="sp_split_particle")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), value :: split_ratio, pmin_split
        end function sp_split_particle

        integer(c_int) function sp_step_once(ctx, n, t0, dtf, state, rands, fixed_dt, deltas) &
                bind(C, name="sp_step_once")
            import :: c_int, c_int8_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n
            real(c_double), value :: t0, dtf
            real(c_double), intent(inout) :: state(*)
            real(c_double), intent(in) :: rands(*)
            integer(c_int8_t), intent(in) :: fixed_dt(*)
            real(c_double), intent(out) :: deltas(*)
        end function sp_step_once

        integer(c_int) function sp_distributions_init(ctx, dc) &
                bind(C, name="sp_distributions_init")
            import :: c_int, c_ptr, sp_dist_config
            type(c_ptr), value :: ctx
            type(sp_dist_config), intent(in) :: dc
        end function sp_distributions_init

        integer(c_int) function sp_distributions(ctx, fglobal, flocal1, flocal2, flocal3, &
                flocal4) bind(C, name="sp_distributions")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, fglobal, flocal1, flocal2, flocal3, flocal4
        end function sp_distributions

        integer(c_int) function sp_local_dist_dims(ctx, set, dims) &
                bind(C, name="sp_local_dist_dims")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: set
            integer(c_int64_t), intent(out) :: dims(5)
        end function sp_local_dist_dims

        integer(c_int) function sp_escaped_distributions(ctx, fescaped, fesc1, fesc2, fesc3, &
                fesc4) bind(C, name="sp_escaped_distributions")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, fescaped, fesc1, fesc2, fesc3, fesc4
        end function sp_escaped_distributions

        integer(c_int) function sp_quick_check(ctx, out) bind(C, name="sp_quick_check")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(9)
        end function sp_quick_check

        integer(c_int) function sp_calibration_copy(ctx, nbytes, ms) &
                bind(C, name="sp_calibration_copy")
            import :: c_int, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: nbytes
            real(c_double), intent(out) :: ms
        end function sp_calibration_copy

        integer(c_int) function sp_autotune_mover(ctx, t0, dtf, num_fine_steps, &
                nsteps_interval, res) bind(C, name="sp_autotune_mover")
            import :: c_int, c_int32_t, c_double, c_ptr, sp_tune_result
            type(c_ptr), value :: ctx
            real(c_double), value :: t0, dtf
            integer(c_int32_t), value :: num_fine_steps, nsteps_interval
            type(sp_tune_result), intent(out) :: res
        end function sp_autotune_mover

        integer(c_int) function sp_set_variants(ctx, nvar, variants) &
                bind(C, name="sp_set_variants")
            import :: c_int, c_int32_t, c_ptr, sp_variant
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nvar
            type(sp_variant), intent(in) :: variants(*)
        end function sp_set_variants

        integer(c_int) function sp_inject_variants(ctx, nvar, nptl, dt, dist_flag, particle_v0, &
                t0, dtf, part_boxes, power_index) bind(C, name="sp_inject_variants")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nvar
            integer(c_int64_t), intent(in) :: nptl(*)
            real(c_double), value :: dt
            integer(c_int32_t), value :: dist_flag
            real(c_double), value :: particle_v0, t0, dtf
            real(c_double), intent(in) :: part_boxes(*)
            real(c_double), value :: power_index
        end function sp_inject_variants

        integer(c_int) function sp_variant_distributions(ctx, nvar, fglobal, counts) &
                bind(C, name="sp_variant_distributions")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nvar
            real(c_double), intent(out) :: fglobal(*)
            integer(c_int64_t), intent(out) :: counts(*)
        end function sp_variant_distributions

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function c_strlen
    end interface

contains

    !< The library's last error message as a Fortran string.
    function sp_error_message() result(msg)
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i
        p = sp_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        enddo
    end function sp_error_message
end module sp_hip_binding
