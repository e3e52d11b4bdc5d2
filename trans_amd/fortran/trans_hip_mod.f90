! iso_c_binding interface to libtrans_hip.so (include/trans_hip.h) for a Fortran host.
!
! The reference's driver calls parameterless routines on module arrays, one block at a
! time (main.F90:65-197).  With this module the same driver keeps its arrays and hands them
! over as views: trans_view() describes an array A(ncomp, lo1:hi1, lo2:hi2, lo3:hi3) (or a
! scalar field A(lo1:hi1, ...)) in place; nothing is copied until trans_hip_set_* runs.
! Derived types mirror the C structs field for field (bind(C) gives the C layout).
module trans_hip_mod
    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: TRANS_OK = 0, TRANS_EINVAL = -1, TRANS_EHIP = -2, &
        TRANS_ENOMEM = -3, TRANS_EUNSUPPORTED = -4, TRANS_ENONFINITE = 1
    integer(c_int), parameter, public :: TRANS_NSTAT = 23
    integer(c_int), parameter, public :: TRANS_FIELD_D = 0, TRANS_FIELD_RDS = 1, TRANS_FIELD_DTM = 2, &
        TRANS_FIELD_DQDXYZ = 3, TRANS_FIELD_SHOCK = 4, TRANS_FIELD_FSCHEME = 5

    type, bind(C), public :: trans_params
        integer(c_int) :: kind_spatial, kind_roe, if_euler, if_turb_org, kind_model, kind_dual, kchld, &
            if_rk, kitm, kind_hybrid, kind_subhy, if_comppress, kind_cc, kind_rotmod, if_precondition, &
            cfl_sub_step, ki_ini, implicit_fp32
        real(c_double) :: xm, rxm2, ppf, csthlnd, csth1, ref, omega1, tiny, vxf, vyf, vzf
        real(c_double) :: dt_lusgs_main, cfl_sub_start, cfl_sub_end, gama_max, gama_min
        real(c_double) :: prod_lmt, visratio, fsti, a1, akapa, beta_star
        real(c_double) :: sigk1, sigo1, beta1, gama1, sigk2, sigo2, beta2, gama2
        real(c_double) :: p0in, t0in, pouthub, ps00, ts00, rkf, rof, rmiuf
        real(c_double) :: f_limiter1, f_limiter2, alscale
    end type trans_params

    ! ConnectivityStruct fields (Module_Block.F90:4-23); target_block/target_conn 0-based
    type, bind(C), public :: trans_patch
        integer(c_int) :: bctype, subbctype, pertype
        integer(c_int) :: point_start(3), point_end(3)
        integer(c_int) :: ijkdirect, left_or_right
        integer(c_int) :: ijkadd(3), cijkadd(3), cijkadd1(3), cijkadd2(3)
        integer(c_int) :: ceijkadd(3), ceijkadd1(3), ceijkadd2(3)
        integer(c_int) :: target_block, target_conn
        integer(c_int) :: transform(3), transform_t(3), rev_matrix(3)
        integer(c_int) :: lcross
        real(c_double) :: rotate_angle
    end type trans_patch

    type, bind(C), public :: trans_block_desc
        integer(c_int) :: ni, nj, nk, npatch
        type(c_ptr) :: patches = c_null_ptr
    end type trans_block_desc

    type, bind(C), public :: trans_fview
        type(c_ptr) :: p = c_null_ptr
        integer(c_int) :: ncomp = 0
        integer(c_int) :: lo(3) = 0, hi(3) = -1
    end type trans_fview

    type, bind(C), public :: trans_geom_view
        type(trans_fview) :: sd, grad, alagm, vol, rad, dst, aleng, aseng, yc, zc, thtc
        type(c_ptr) :: markwall(6) = c_null_ptr
    end type trans_geom_view

    type, bind(C), public :: trans_plane_integrals
        real(c_double) :: mass_flow, area, p0_mass_avg, t0_mass_avg, p0_abs_mass_avg, t0_abs_mass_avg
    end type trans_plane_integrals

    type, bind(C), public :: trans_state_view
        type(trans_fview) :: v, pp, t, rmiu, wt0, wt1, wt2
    end type trans_state_view

    type, bind(C), public :: trans_resid
        real(c_double) :: drsa, dua
        integer(c_int) :: ima, jma, kma
        real(c_double) :: drsk, duk
        integer(c_int) :: imk, jmk, kmk
        real(c_double) :: drso, duo
        integer(c_int) :: imo, jmo, kmo
    end type trans_resid

    ! serving (trans_srv_* in include/trans_hip.h)
    integer(c_int), parameter, public :: TRANS_SRV_QUEUED = 0, TRANS_SRV_RUNNING = 1, &
        TRANS_SRV_CONVERGED = 2, TRANS_SRV_MAX_ITER = 3, TRANS_SRV_CANCELLED = 4, TRANS_SRV_FAILED = 5

    type, bind(C), public :: trans_srv_config
        integer(c_size_t) :: hbm_budget
        integer(c_int) :: max_active, chunk
    end type trans_srv_config

    type, bind(C), public :: trans_srv_request
        type(trans_params) :: params
        integer(c_int) :: nblk
        type(c_ptr) :: blocks = c_null_ptr, geom = c_null_ptr, state = c_null_ptr
        real(c_double) :: target_drop
        integer(c_int) :: max_iter
    end type trans_srv_request

    type, bind(C), public :: trans_srv_status
        integer(c_int) :: state, error, iterations, iterations_to_target
        real(c_double) :: drsa_first, drsa_last, t_submit, t_admit, t_done
        integer(c_size_t) :: footprint
    end type trans_srv_status

    ! outputs of trans_geom_metrics (c_loc of each -2:N+2 array; sd, grad, alagm component first)
    type, bind(C), public :: trans_geom_metrics_out
        type(c_ptr) :: xc, yc, zc, thtc, rad, aleng, aseng, vol, sd, grad, alagm
    end type trans_geom_metrics_out

    public :: trans_view, trans_error_text

    interface
        integer(c_int) function trans_hip_create(params, nblk, blocks, ctx) bind(C, name="trans_hip_create")
            import :: c_int, c_ptr, trans_params, trans_block_desc
            type(trans_params), intent(in) :: params
            integer(c_int), value :: nblk
            type(trans_block_desc), intent(in) :: blocks(*)
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function trans_hip_destroy(ctx) bind(C, name="trans_hip_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function trans_hip_last_error(ctx) bind(C, name="trans_hip_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function trans_hip_set_params(ctx, params) bind(C, name="trans_hip_set_params")
            import :: c_int, c_ptr, trans_params
            type(c_ptr), value :: ctx
            type(trans_params), intent(in) :: params
        end function
        integer(c_int) function trans_hip_set_iteration(ctx, ki) bind(C, name="trans_hip_set_iteration")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: ki
        end function
        integer(c_int) function trans_hip_set_geometry(ctx, blk, g) bind(C, name="trans_hip_set_geometry")
            import :: c_int, c_ptr, trans_geom_view
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk
            type(trans_geom_view), intent(in) :: g
        end function
        integer(c_int) function trans_hip_set_state(ctx, blk, s) bind(C, name="trans_hip_set_state")
            import :: c_int, c_ptr, trans_state_view
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk
            type(trans_state_view), intent(in) :: s
        end function
        integer(c_int) function trans_hip_get_state(ctx, blk, s) bind(C, name="trans_hip_get_state")
            import :: c_int, c_ptr, trans_state_view
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk
            type(trans_state_view), intent(in) :: s
        end function
        integer(c_int) function trans_hip_get_field(ctx, blk, which, out) bind(C, name="trans_hip_get_field")
            import :: c_int, c_ptr, trans_fview
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk, which
            type(trans_fview), intent(in) :: out
        end function
        integer(c_int) function trans_hip_spectral_radius(ctx) bind(C, name="trans_hip_spectral_radius")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function trans_hip_bc(ctx, sweep) bind(C, name="trans_hip_bc")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: sweep
        end function
        integer(c_int) function trans_hip_space(ctx) bind(C, name="trans_hip_space")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function trans_hip_residual(ctx) bind(C, name="trans_hip_residual")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function trans_hip_lusgs_update(ctx, per_block) bind(C, name="trans_hip_lusgs_update")
            import :: c_int, c_ptr, trans_resid
            type(c_ptr), value :: ctx
            type(trans_resid), intent(out) :: per_block(*)
        end function
        integer(c_int) function trans_hip_rk_update(ctx, itr_chld, per_block) bind(C, name="trans_hip_rk_update")
            import :: c_int, c_ptr, trans_resid
            type(c_ptr), value :: ctx
            integer(c_int), value :: itr_chld
            type(trans_resid), intent(out) :: per_block(*)
        end function
        integer(c_int) function trans_hip_update_u(ctx) bind(C, name="trans_hip_update_u")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function trans_hip_iterate(ctx, n_iter, per_block) bind(C, name="trans_hip_iterate")
            import :: c_int, c_ptr, trans_resid
            type(c_ptr), value :: ctx
            integer(c_int), value :: n_iter
            type(trans_resid), intent(out) :: per_block(*)
        end function
        integer(c_int) function trans_hip_plane_integrals(ctx, blk, dir, index, out) &
                bind(C, name="trans_hip_plane_integrals")
            import :: c_int, c_ptr, trans_plane_integrals
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk, dir, index
            type(trans_plane_integrals), intent(out) :: out
        end function
        integer(c_int) function trans_hip_sync(ctx) bind(C, name="trans_hip_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        ! path arguments: null-terminated, e.g. trim(ThisBlock%FLDAT)//c_null_char
        integer(c_int) function trans_hip_write_restart(ctx, blk, ki, path) bind(C, name="trans_hip_write_restart")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk, ki
            character(kind=c_char), intent(in) :: path(*)
        end function
        integer(c_int) function trans_hip_read_restart(ctx, blk, path, ki) bind(C, name="trans_hip_read_restart")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), intent(out) :: ki
        end function
        integer(c_int) function trans_hip_stats_enable(ctx, ki_aver, ki_statis) bind(C, name="trans_hip_stats_enable")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: ki_aver, ki_statis
        end function
        integer(c_int) function trans_hip_average_fields(ctx) bind(C, name="trans_hip_average_fields")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function trans_hip_get_stats(ctx, blk, out) bind(C, name="trans_hip_get_stats")
            import :: c_int, c_ptr, trans_fview
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk
            type(trans_fview), intent(in) :: out
        end function
        integer(c_int) function trans_hip_set_stats(ctx, blk, in) bind(C, name="trans_hip_set_stats")
            import :: c_int, c_ptr, trans_fview
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk
            type(trans_fview), intent(in) :: in
        end function
        integer(c_int) function trans_hip_write_stats(ctx, blk, xc, yc, zc, path) bind(C, name="trans_hip_write_stats")
            import :: c_int, c_ptr, c_char, trans_fview
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk
            type(trans_fview), intent(in) :: xc, yc, zc
            character(kind=c_char), intent(in) :: path(*)
        end function
        ! Tecplot ASCII instantaneous field of block blk (0-based): point coordinates XX, YY, ZZ
        ! and the 31 cell-centred variables of Output/OUTPUT_InstField.F90; scales = (/Roref, Vref,
        ! Pref, Tinf/)
        integer(c_int) function trans_hip_write_tecplot(ctx, blk, x, y, z, scales, path) &
                bind(C, name="trans_hip_write_tecplot")
            import :: c_int, c_ptr, c_char, c_double, trans_fview
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk
            type(trans_fview), intent(in) :: x, y, z
            real(c_double), intent(in) :: scales(4)
            character(kind=c_char), intent(in) :: path(*)
        end function
        integer(c_int) function trans_hip_read_stats(ctx, blk, path) bind(C, name="trans_hip_read_stats")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ctx
            integer(c_int), value :: blk
            character(kind=c_char), intent(in) :: path(*)
        end function
        integer(c_int) function trans_hip_footprint(params, nblk, blocks, bytes) bind(C, name="trans_hip_footprint")
            import :: c_int, c_size_t, trans_params, trans_block_desc
            type(trans_params), intent(in) :: params
            integer(c_int), value :: nblk
            type(trans_block_desc), intent(in) :: blocks(*)
            integer(c_size_t), intent(out) :: bytes
        end function
        integer(c_int) function trans_srv_create(cfg, srv) bind(C, name="trans_srv_create")
            import :: c_int, c_ptr, trans_srv_config
            type(trans_srv_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: srv
        end function
        integer(c_int) function trans_srv_submit(srv, req, id) bind(C, name="trans_srv_submit")
            import :: c_int, c_ptr, trans_srv_request
            type(c_ptr), value :: srv
            type(trans_srv_request), intent(in) :: req
            integer(c_int), intent(out) :: id
        end function
        integer(c_int) function trans_srv_progress(srv, wait, n_unfinished) bind(C, name="trans_srv_progress")
            import :: c_int, c_ptr
            type(c_ptr), value :: srv
            integer(c_int), value :: wait
            integer(c_int), intent(out) :: n_unfinished
        end function
        integer(c_int) function trans_srv_poll(srv, id, out) bind(C, name="trans_srv_poll")
            import :: c_int, c_ptr, trans_srv_status
            type(c_ptr), value :: srv
            integer(c_int), value :: id
            type(trans_srv_status), intent(out) :: out
        end function
        integer(c_int) function trans_srv_cancel(srv, id) bind(C, name="trans_srv_cancel")
            import :: c_int, c_ptr
            type(c_ptr), value :: srv
            integer(c_int), value :: id
        end function
        integer(c_int) function trans_srv_result(srv, id, blk, out) bind(C, name="trans_srv_result")
            import :: c_int, c_ptr, trans_state_view
            type(c_ptr), value :: srv
            integer(c_int), value :: id, blk
            type(trans_state_view), intent(in) :: out
        end function
        integer(c_int) function trans_srv_release(srv, id) bind(C, name="trans_srv_release")
            import :: c_int, c_ptr
            type(c_ptr), value :: srv
            integer(c_int), value :: id
        end function
        integer(c_int) function trans_srv_destroy(srv) bind(C, name="trans_srv_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: srv
        end function
        integer(c_int) function trans_geom_wall_distance(cells, ncell, panels, npanel, dst) &
                bind(C, name="trans_geom_wall_distance")
            import :: c_int, c_long, c_double
            real(c_double), intent(in) :: cells(*), panels(*)
            integer(c_long), value :: ncell, npanel
            real(c_double), intent(out) :: dst(*)
        end function
        integer(c_int) function trans_geom_metrics(ni, nj, nk, xx, yy, zz, tiny, out) &
                bind(C, name="trans_geom_metrics")
            import :: c_int, c_double, trans_geom_metrics_out
            integer(c_int), value :: ni, nj, nk
            real(c_double), intent(in) :: xx(*), yy(*), zz(*)
            real(c_double), value :: tiny
            type(trans_geom_metrics_out), intent(in) :: out
        end function
    end interface
    public :: trans_hip_create, trans_hip_destroy, trans_hip_last_error, trans_hip_set_params, &
        trans_hip_set_iteration, trans_hip_set_geometry, trans_hip_set_state, trans_hip_get_state, &
        trans_hip_get_field, trans_hip_spectral_radius, trans_hip_bc, trans_hip_space, &
        trans_hip_residual, trans_hip_lusgs_update, trans_hip_rk_update, trans_hip_update_u, trans_hip_iterate, &
        trans_hip_plane_integrals, trans_hip_sync, trans_hip_footprint, trans_srv_create, trans_srv_submit, &
        trans_hip_write_restart, trans_hip_read_restart, trans_hip_stats_enable, trans_hip_average_fields, &
        trans_hip_get_stats, trans_hip_set_stats, trans_hip_write_stats, trans_hip_read_stats, trans_hip_write_tecplot, &
        trans_srv_progress, trans_srv_poll, trans_srv_cancel, trans_srv_result, trans_srv_release, trans_srv_destroy, &
        trans_geom_wall_distance, trans_geom_metrics

contains

    ! View of the contiguous array whose first element is `first`, with ncomp components
    ! and point bounds lo(3):hi(3) (e.g. -2:N+2 for V, Vol, SD; 1:N+2 for Wt0).
    function trans_view(first, ncomp, lo, hi) result(v)
        real(c_double), intent(in), target :: first
        integer, intent(in) :: ncomp, lo(3), hi(3)
        type(trans_fview) :: v
        v%p = c_loc(first)
        v%ncomp = ncomp
        v%lo = lo
        v%hi = hi
    end function trans_view

    ! Text of the context's last error.
    function trans_error_text(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=256) :: msg
        character(kind=c_char), pointer :: s(:)
        integer :: n
        msg = ' '
        call c_f_pointer(trans_hip_last_error(ctx), s, [256])
        do n = 1, 256
            if (s(n) == c_null_char) exit
            msg(n:n) = s(n)
        end do
    end function trans_error_text
end module trans_hip_mod
