! Fortran host driving libtrans_hip.so through trans_hip_mod, the way the reference's
! main.F90 would: the host owns its arrays in the reference's layout (component first,
! bounds -2:N+2), describes them as views and calls one library function per routine.
!
! Case: a Cartesian box in inviscid flow at M = 0.5 along x, far field at imin/imax,
! symmetry on the other faces (connectivity rows as PLOT3D_Read_And_Init builds them,
! Input/PLOT3DReadInit.F90:183-348).  Checks:
!   1. uniform flow stays uniform (max |V - V0| <= 1e-12, DRSa <= 1e-13);
!   2. a +0.1 % density bump in one cell is convected and damped: the norms stay finite
!      and the disturbance decreases.
! Prints one line "box_demo ok ..." and exits 0, or prints the failure and stops with 1.
program box_demo
    use iso_c_binding
    use trans_hip_mod
    implicit none
    integer, parameter :: ni = 33, nj = 9, nk = 5
    real(c_double), allocatable, target :: V(:, :, :, :), PP(:, :, :), T(:, :, :), Rmiu(:, :, :)
    real(c_double), allocatable, target :: Wt0(:, :, :, :), Wt1(:, :, :, :)
    real(c_double), allocatable, target :: SD(:, :, :, :), Grad(:, :, :, :), Alagm(:, :, :, :)
    real(c_double), allocatable, target :: Vol(:, :, :), rad(:, :, :), Dst(:, :, :), aLeng(:, :, :)
    real(c_double), allocatable, target :: Yc(:, :, :), Zc(:, :, :), thtc(:, :, :)
    real(c_double), allocatable :: V0(:, :, :, :)
    type(trans_patch), target :: pt(6)
    type(trans_block_desc) :: bd(1)
    type(trans_params) :: prm
    type(trans_geom_view) :: gv
    type(trans_state_view) :: sv
    type(trans_resid) :: res(1)
    type(c_ptr) :: ctx
    integer :: lo(3), hi(3), dims(3), f, d, rc, j, k, it
    real(c_double) :: dx, dy, dz, xm, p_inf, dev0, dev1, r0

    xm = 0.5d0
    dx = 1.0d0 / (ni - 1); dy = 0.25d0 / (nj - 1); dz = 0.1d0 / (nk - 1)
    lo = [-2, -2, -2]; hi = [ni + 2, nj + 2, nk + 2]; dims = [ni, nj, nk]
    allocate (V(7, -2:ni + 2, -2:nj + 2, -2:nk + 2), Wt0(7, -2:ni + 2, -2:nj + 2, -2:nk + 2))
    allocate (Wt1(7, -2:ni + 2, -2:nj + 2, -2:nk + 2), V0(7, -2:ni + 2, -2:nj + 2, -2:nk + 2))
    allocate (PP(-2:ni + 2, -2:nj + 2, -2:nk + 2), T(-2:ni + 2, -2:nj + 2, -2:nk + 2))
    allocate (Rmiu(-2:ni + 2, -2:nj + 2, -2:nk + 2))
    allocate (SD(9, -2:ni + 2, -2:nj + 2, -2:nk + 2), Grad(3, -2:ni + 2, -2:nj + 2, -2:nk + 2))
    allocate (Alagm(3, -2:ni + 2, -2:nj + 2, -2:nk + 2))
    allocate (Vol(-2:ni + 2, -2:nj + 2, -2:nk + 2), rad(-2:ni + 2, -2:nj + 2, -2:nk + 2))
    allocate (Dst(-2:ni + 2, -2:nj + 2, -2:nk + 2), aLeng(-2:ni + 2, -2:nj + 2, -2:nk + 2))
    allocate (Yc(-2:ni + 2, -2:nj + 2, -2:nk + 2), Zc(-2:ni + 2, -2:nj + 2, -2:nk + 2))
    allocate (thtc(-2:ni + 2, -2:nj + 2, -2:nk + 2))

    ! geometry of a Cartesian box (Geom/CalcDst.F90 reduces to these on axis-aligned cells)
    SD = 0.0d0
    SD(1, :, :, :) = dy * dz; SD(5, :, :, :) = dx * dz; SD(9, :, :, :) = dx * dy
    Grad(1, :, :, :) = dy * dz; Grad(2, :, :, :) = dx * dz; Grad(3, :, :, :) = dx * dy
    Alagm(1, :, :, :) = dx; Alagm(2, :, :, :) = dy; Alagm(3, :, :, :) = dz
    Vol = dx * dy * dz
    do k = -2, nk + 2
        do j = -2, nj + 2
            Yc(:, j, k) = 1.0d0 + (j - 0.5d0) * dy
            Zc(:, j, k) = (k - 0.5d0) * dz
        end do
    end do
    rad = sqrt(Yc**2 + Zc**2)
    thtc = atan2(Zc, Yc)
    Dst = 1.0d3
    aLeng = max(dx, dy, dz)

    ! deck: Euler, MUSCL + Roe, LU-SGS CFL 5, steady dual time (Input/ParameterSet.F90)
    prm%kind_spatial = 2; prm%kind_roe = 1; prm%if_euler = 1; prm%if_turb_org = 0
    prm%kind_model = 1; prm%kind_dual = 2; prm%kchld = 1; prm%if_rk = 0; prm%kitm = 1
    prm%kind_hybrid = 0; prm%kind_subhy = 0; prm%if_comppress = 0; prm%kind_cc = 0
    prm%kind_rotmod = 0; prm%if_precondition = 0; prm%cfl_sub_step = 1; prm%ki_ini = 0
    prm%implicit_fp32 = 0
    prm%xm = xm; prm%rxm2 = 1.4d0 * xm * xm; prm%ppf = 1.0d0 / (1.4d0 * xm * xm)
    prm%csthlnd = 110.0d0 / 288.0d0; prm%csth1 = 288.0d0 / 273.15d0; prm%ref = 1.0d6
    prm%omega1 = 0.0d0; prm%tiny = 1.0d-20; prm%vxf = 1.0d0; prm%vyf = 0.0d0; prm%vzf = 0.0d0
    prm%dt_lusgs_main = 1.0d10; prm%cfl_sub_start = 5.0d0; prm%cfl_sub_end = 5.0d0
    prm%gama_max = 0.012d0; prm%gama_min = 0.001d0; prm%prod_lmt = 20.0d0; prm%visratio = 1.0d0
    prm%fsti = 0.01d0; prm%a1 = 0.31d0; prm%akapa = 0.41d0; prm%beta_star = 0.09d0
    prm%sigk1 = 0.85d0; prm%sigo1 = 0.5d0; prm%beta1 = 0.075d0; prm%gama1 = 5.0d0 / 9.0d0
    prm%sigk2 = 1.0d0; prm%sigo2 = 0.856d0; prm%beta2 = 0.0828d0; prm%gama2 = 0.44d0
    prm%p0in = 0.0d0; prm%t0in = 1.0d0; prm%pouthub = 0.0d0; prm%ps00 = 0.0d0; prm%ts00 = 1.0d0
    prm%rkf = 1.5d-4; prm%rof = 1.2d0 * 1.5d-4 / 1.5d-5; prm%rmiuf = 1.5d-5
    prm%f_limiter1 = 0.0d0; prm%f_limiter2 = 0.0d0; prm%alscale = 1.0d0

    ! uniform freestream: rho = 1, u = 1, p = 1/(gamma M^2), T = gamma M^2 p / rho = 1
    p_inf = prm%ppf
    V = 0.0d0
    V(1, :, :, :) = 1.0d0; V(2, :, :, :) = 1.0d0; V(5, :, :, :) = p_inf / 0.4d0 + 0.5d0
    PP = p_inf; T = prm%rxm2 * p_inf; Rmiu = 0.0d0
    Wt0 = V; Wt1 = V

    ! connectivity rows: imin, imax far field; jmin, jmax, kmin, kmax symmetry
    do f = 1, 6
        d = (f + 1) / 2
        pt(f) = empty_patch()
        pt(f)%bctype = merge(2, 3, d == 1)
        pt(f)%point_start = 1
        pt(f)%point_end = [ni - 1, nj - 1, nk - 1]
        pt(f)%ijkdirect = d
        pt(f)%transform(d) = 1
        if (mod(f, 2) == 1) then  ! min face
            pt(f)%left_or_right = 1
            pt(f)%point_end(d) = 1
            pt(f)%ijkadd(d) = 1; pt(f)%cijkadd(d) = 0; pt(f)%ceijkadd(d) = -1
            pt(f)%cijkadd1(d) = 1; pt(f)%cijkadd2(d) = 2; pt(f)%ceijkadd1(d) = -2; pt(f)%ceijkadd2(d) = -3
        else                      ! max face
            pt(f)%left_or_right = 2
            pt(f)%point_start(d) = dims(d)
            pt(f)%point_end(d) = dims(d)
            pt(f)%ijkadd(d) = -1; pt(f)%cijkadd(d) = -1; pt(f)%ceijkadd(d) = 0
            pt(f)%cijkadd1(d) = -2; pt(f)%cijkadd2(d) = -3; pt(f)%ceijkadd1(d) = 1; pt(f)%ceijkadd2(d) = 2
        end if
    end do
    bd(1)%ni = ni; bd(1)%nj = nj; bd(1)%nk = nk; bd(1)%npatch = 6
    bd(1)%patches = c_loc(pt(1))

    call check(trans_hip_create(prm, 1_c_int, bd, ctx), "create")
    gv%sd = trans_view(SD(1, -2, -2, -2), 9, lo, hi)
    gv%grad = trans_view(Grad(1, -2, -2, -2), 3, lo, hi)
    gv%alagm = trans_view(Alagm(1, -2, -2, -2), 3, lo, hi)
    gv%vol = trans_view(Vol(-2, -2, -2), 1, lo, hi)
    gv%rad = trans_view(rad(-2, -2, -2), 1, lo, hi)
    gv%dst = trans_view(Dst(-2, -2, -2), 1, lo, hi)
    gv%aleng = trans_view(aLeng(-2, -2, -2), 1, lo, hi)
    gv%yc = trans_view(Yc(-2, -2, -2), 1, lo, hi)
    gv%zc = trans_view(Zc(-2, -2, -2), 1, lo, hi)
    gv%thtc = trans_view(thtc(-2, -2, -2), 1, lo, hi)
    call check(trans_hip_set_geometry(ctx, 0_c_int, gv), "set_geometry")
    sv%v = trans_view(V(1, -2, -2, -2), 7, lo, hi)
    sv%pp = trans_view(PP(-2, -2, -2), 1, lo, hi)
    sv%t = trans_view(T(-2, -2, -2), 1, lo, hi)
    sv%rmiu = trans_view(Rmiu(-2, -2, -2), 1, lo, hi)
    sv%wt0 = trans_view(Wt0(1, -2, -2, -2), 7, lo, hi)
    sv%wt1 = trans_view(Wt1(1, -2, -2, -2), 7, lo, hi)
    call check(trans_hip_set_state(ctx, 0_c_int, sv), "set_state")
    call check(trans_hip_set_iteration(ctx, 1_c_int), "set_iteration")

    ! 1. freestream preservation
    V0 = V
    call check(trans_hip_iterate(ctx, 5_c_int, res), "iterate")
    call check(trans_hip_get_state(ctx, 0_c_int, sv), "get_state")
    dev0 = maxval(abs(V(:, 1:ni - 1, 1:nj - 1, 1:nk - 1) - V0(:, 1:ni - 1, 1:nj - 1, 1:nk - 1)))
    if (dev0 > 1.0d-12 .or. res(1)%drsa > 1.0d-13) then
        print '(a, es12.4, a, es12.4)', "box_demo FAILED: freestream drift ", dev0, " DRSa ", res(1)%drsa
        stop 1
    end if

    ! 2. a density bump (pressure kept, so T follows) is convected out and damped
    V(1, 8, 4, 2) = V(1, 8, 4, 2) * 1.001d0
    V(2, 8, 4, 2) = V(1, 8, 4, 2)
    V(5, 8, 4, 2) = p_inf / 0.4d0 + 0.5d0 * V(1, 8, 4, 2)
    T(8, 4, 2) = prm%rxm2 * p_inf / V(1, 8, 4, 2)
    call check(trans_hip_set_state(ctx, 0_c_int, sv), "set_state")
    dev0 = maxval(abs(V(1, 1:ni - 1, 1:nj - 1, 1:nk - 1) - 1.0d0))
    call check(trans_hip_iterate(ctx, 1_c_int, res), "iterate")
    r0 = res(1)%drsa
    do it = 1, 4
        call check(trans_hip_iterate(ctx, 20_c_int, res), "iterate")
    end do
    call check(trans_hip_get_state(ctx, 0_c_int, sv), "get_state")
    dev1 = maxval(abs(V(1, 1:ni - 1, 1:nj - 1, 1:nk - 1) - 1.0d0))
    if (.not. (res(1)%drsa == res(1)%drsa) .or. .not. (dev1 < dev0) .or. .not. (res(1)%drsa < r0)) then
        print '(a, 4es12.4)', "box_demo FAILED: disturbance not damped ", dev0, dev1, r0, res(1)%drsa
        stop 1
    end if
    rc = trans_hip_destroy(ctx)
    print '(a, es11.4, a, es11.4, a, es11.4, a, es11.4)', "box_demo ok: DRSa ", r0, " -> ", res(1)%drsa, &
        ", max|rho-1| ", dev0, " -> ", dev1

contains

    function empty_patch() result(p)
        type(trans_patch) :: p
        p%bctype = 0; p%subbctype = 0; p%pertype = 0
        p%point_start = 0; p%point_end = 0; p%ijkdirect = 0; p%left_or_right = 0
        p%ijkadd = 0; p%cijkadd = 0; p%cijkadd1 = 0; p%cijkadd2 = 0
        p%ceijkadd = 0; p%ceijkadd1 = 0; p%ceijkadd2 = 0
        p%target_block = 0; p%target_conn = 0
        p%transform = 0; p%transform_t = 0; p%rev_matrix = 0; p%lcross = 0
        p%rotate_angle = 0.0d0
    end function empty_patch

    subroutine check(code, what)
        integer(c_int), intent(in) :: code
        character(len=*), intent(in) :: what
        if (code /= TRANS_OK) then
            print '(a, a, a, i0, a, a)', "box_demo FAILED: ", what, " returned ", code, ": ", &
                trim(trans_error_text(ctx))
            stop 1
        end if
    end subroutine check
end program box_demo
