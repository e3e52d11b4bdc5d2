! TEST INFRASTRUCTURE ONLY.  C-callable wrappers over routines of the reference that compile
! unmodified from its own sources: the low-Mach preconditioning (getPrecondPara, PrecondDW;
! Common/UTIL_Precond_Funcs.F90:7-31, 91-129), the MUSCL limiters (ALIMITER, MUSCL2;
! Space/UTIL_MUSCL.F90), the AUSMPW+ face flux (CONVECT_AUSMPW_Plus_Common;
! Space/CONVECT_AUSMPW_Plus.F90:204-291), the hybrid RANS/LES functions (FSST,
! fd_function, fe_function; Time/Hybrid_function.F90), and whole-block routines run on a
! block whose arrays are the oracle's (the convective reconstructions and fluxes of Space/,
! TIME_Update_U and the Runge-Kutta stage routines).  oracle/Makefile (target ref) compiles
! those files where they lie, with default REAL promoted to 8 bytes as the FP64 build of the
! reference is, and links them with this driver into oracle/_ref/libtrans_ref.so.  Only tests/
! load that library, to check the oracle's restatements against the reference bit for bit.
!
! The routines read module variables of Module_Global.F90 (PPf, Tiny, XM, Csthlnd, Csth1, Ref,
! Kind_SubHY, gama_max, gama_min, F_limiter2, KI, Itr_Chld, Kind_Dual, Kind_Model, A1,
! Beta_Star, Prod_Lmt, RXM2, If_Euler, IF_Turb, KItm); each
! wrapper sets the ones its routines read before the calls.

subroutine ref_precond_para(xm_in, csthlnd_in, csth1_in, ref_in, vm2, am2, rr, tt, rmiut, alagmr, ubetac2) &
    bind(C, name="ref_precond_para")
    use iso_c_binding, only: c_double
    use Global, only: XM, Csthlnd, Csth1, Ref
    implicit none
    real(c_double), value :: xm_in, csthlnd_in, csth1_in, ref_in
    real(c_double), value :: vm2, am2, rr, tt, rmiut, alagmr
    real(c_double), intent(out) :: ubetac2
    real(8) :: ub
    XM = xm_in
    Csthlnd = csthlnd_in
    Csth1 = csth1_in
    Ref = ref_in
    call getPrecondPara(vm2, am2, rr, tt, rmiut, alagmr, ub)
    ubetac2 = ub
end subroutine

subroutine ref_precond_dw(ppf_in, a, pro, pvx, pvy, pvz, pe, aac, ubetac2, ppp, ptt) bind(C, name="ref_precond_dw")
    use iso_c_binding, only: c_double
    use Global, only: PPf
    implicit none
    real(c_double), value :: ppf_in
    real(c_double), intent(inout) :: a(5)
    real(c_double), value :: pro, pvx, pvy, pvz, pe, aac, ubetac2, ppp, ptt
    real(8) :: ro, rx, ry, rz, re
    PPf = ppf_in
    ro = a(1)
    rx = a(2)
    ry = a(3)
    rz = a(4)
    re = a(5)
    call PrecondDW(ro, rx, ry, rz, re, pro, pvx, pvy, pvz, pe, aac, ubetac2, ppp, ptt)
    a(1) = ro
    a(2) = rx
    a(3) = ry
    a(4) = rz
    a(5) = re
end subroutine

! ALIMITER and MUSCL2 (Space/UTIL_MUSCL.F90:2-34): the MUSCL slope limiters.
subroutine ref_limiters(v1, v2, alim, mus2) bind(C, name="ref_limiters")
    use iso_c_binding, only: c_double
    implicit none
    real(c_double), value :: v1, v2
    real(c_double), intent(out) :: alim, mus2
    real(8), external :: ALIMITER, MUSCL2
    real(8) :: a, b
    a = v1
    b = v2
    alim = ALIMITER(a, b)
    mus2 = MUSCL2(a, b)
end subroutine

! CONVECT_AUSMPW_Plus_Common (Space/CONVECT_AUSMPW_Plus.F90:204-291): the AUSMPW+ face flux
! from the left and right states; it reads PPf and Tiny.
subroutine ref_ausmpw_common(ppf_in, tiny_in, pgam, s, rmiut, alagmr, out) bind(C, name="ref_ausmpw_common")
    use iso_c_binding, only: c_double
    use Global, only: PPf, Tiny
    implicit none
    real(c_double), value :: ppf_in, tiny_in, pgam, rmiut, alagmr
    real(c_double), intent(in) :: s(14)
    real(c_double), intent(out) :: out(5)
    real(8) :: g, x(14), mt, al, r(5)
    PPf = ppf_in
    Tiny = tiny_in
    g = pgam
    x = s
    mt = rmiut
    al = alagmr
    call CONVECT_AUSMPW_Plus_Common(g, x(1), x(2), x(3), x(4), x(5), x(6), x(7), x(8), x(9), x(10), &
                                    x(11), x(12), x(13), x(14), mt, al, r(1), r(2), r(3), r(4), r(5))
    out = r
end subroutine

! FSST, fd_function and fe_function (Time/Hybrid_function.F90:1-95): the DDES shielding
! function for Kind_SubHY 0-3 and the IDDES blending functions; they read Kind_SubHY and Ref.
subroutine ref_hybrid(kind_subhy_in, ref_in, den, vk, vomega, dist, rmiul, cdiff, rmiut, aleng, g, res) &
    bind(C, name="ref_hybrid")
    use iso_c_binding, only: c_double, c_int
    use Global, only: Kind_SubHY, Ref
    implicit none
    integer(c_int), value :: kind_subhy_in
    real(c_double), value :: ref_in, den, vk, vomega, dist, rmiul, cdiff, rmiut, aleng
    real(c_double), intent(in) :: g(9)
    real(c_double), intent(out) :: res(3)
    real(8), external :: FSST, fd_function, fe_function
    real(8) :: x(9), d, k, w, ds, ml, cd, mt, al
    Kind_SubHY = kind_subhy_in
    Ref = ref_in
    x = g
    d = den
    k = vk
    w = vomega
    ds = dist
    ml = rmiul
    cd = cdiff
    mt = rmiut
    al = aleng
    res(1) = FSST(d, k, w, ds, ml, cd)
    res(2) = fd_function(d, mt, ds, al, x(1), x(2), x(3), x(4), x(5), x(6), x(7), x(8), x(9))
    res(3) = fe_function(d, mt, ml, ds, al, x(1), x(2), x(3), x(4), x(5), x(6), x(7), x(8), x(9))
end subroutine

! Whole-block routines.  ref_block_dims sets the block's dimensions (points ni, nj, nk) in a
! block of this driver's own and points the module's ThisBlock at it; ref_bind points one of
! the module's block arrays (Module_Global.F90:190-269) at an oracle array, which holds the
! reference's layout (component first, every direction -2:N+2); ref_block_scalars sets the
! module variables the routines read; ref_block_call runs one routine on the bound block.
module ref_block_state
    use Block, only: BlockStruct
    implicit none
    type(BlockStruct), target, save :: blk
end module

subroutine ref_block_dims(ni, nj, nk) bind(C, name="ref_block_dims")
    use iso_c_binding, only: c_int
    use Global, only: ThisBlock
    use ref_block_state, only: blk
    implicit none
    integer(c_int), value :: ni, nj, nk
    blk%NI = ni
    blk%NJ = nj
    blk%NK = nk
    blk%NDim = [ni, nj, nk]
    blk%NI1 = ni - 1
    blk%NJ1 = nj - 1
    blk%NK1 = nk - 1
    blk%NI2 = ni - 2
    blk%NJ2 = nj - 2
    blk%NK2 = nk - 2
    blk%NI3 = ni - 3
    blk%NJ3 = nj - 3
    blk%NK3 = nk - 3
    blk%NIp1 = ni + 1
    blk%NJp1 = nj + 1
    blk%NKp1 = nk + 1
    blk%NIp2 = ni + 2
    blk%NJp2 = nj + 2
    blk%NKp2 = nk + 2
    blk%MI = ni + 2
    blk%MJ = nj + 2
    blk%MK = nk + 2
    ThisBlock => blk
end subroutine

subroutine ref_bind(which, p) bind(C, name="ref_bind")
    use iso_c_binding, only: c_int, c_ptr, c_f_pointer
    use Global, only: ThisBlock, V, PP, T, rad, rccl, rccr, shock, VL, VR, D, SD, Grad, Alagm, &
                      Dtm, Vol, Wt0, Wt1, Wt2, DQdxyz, Rmiu, Dst
    implicit none
    integer(c_int), value :: which
    type(c_ptr), value :: p
    real(8), pointer :: f(:)
    integer :: a, b, c, n
    a = ThisBlock%NI + 2
    b = ThisBlock%NJ + 2
    c = ThisBlock%NK + 2
    n = (a + 3) * (b + 3) * (c + 3)
    select case (which)
    case (1)
        call c_f_pointer(p, f, [7 * n])
        V(1:7, -2:a, -2:b, -2:c) => f
    case (2)
        call c_f_pointer(p, f, [n])
        PP(-2:a, -2:b, -2:c) => f
    case (3)
        call c_f_pointer(p, f, [n])
        T(-2:a, -2:b, -2:c) => f
    case (4)
        call c_f_pointer(p, f, [n])
        rad(-2:a, -2:b, -2:c) => f
    case (5)
        call c_f_pointer(p, f, [3 * n])
        rccl(1:3, -2:a, -2:b, -2:c) => f
    case (6)
        call c_f_pointer(p, f, [3 * n])
        rccr(1:3, -2:a, -2:b, -2:c) => f
    case (7)
        call c_f_pointer(p, f, [n])
        shock(-2:a, -2:b, -2:c) => f
    case (8)
        call c_f_pointer(p, f, [24 * n])
        VL(1:3, 1:8, -2:a, -2:b, -2:c) => f
    case (9)
        call c_f_pointer(p, f, [24 * n])
        VR(1:3, 1:8, -2:a, -2:b, -2:c) => f
    case (10)
        call c_f_pointer(p, f, [7 * n])
        D(1:7, -2:a, -2:b, -2:c) => f
    case (11)
        call c_f_pointer(p, f, [9 * n])
        SD(1:3, 1:3, -2:a, -2:b, -2:c) => f
    case (12)
        call c_f_pointer(p, f, [3 * n])
        Grad(1:3, -2:a, -2:b, -2:c) => f
    case (13)
        call c_f_pointer(p, f, [3 * n])
        Alagm(1:3, -2:a, -2:b, -2:c) => f
    case (14)
        call c_f_pointer(p, f, [n])
        Dtm(-2:a, -2:b, -2:c) => f
    case (15)
        call c_f_pointer(p, f, [n])
        Vol(-2:a, -2:b, -2:c) => f
    case (16)
        call c_f_pointer(p, f, [7 * n])
        Wt0(1:7, -2:a, -2:b, -2:c) => f
    case (17)
        call c_f_pointer(p, f, [7 * n])
        Wt1(1:7, -2:a, -2:b, -2:c) => f
    case (18)
        call c_f_pointer(p, f, [7 * n])
        Wt2(1:7, -2:a, -2:b, -2:c) => f
    case (19)
        call c_f_pointer(p, f, [18 * n])
        DQdxyz(1:18, -2:a, -2:b, -2:c) => f
    case (20)
        call c_f_pointer(p, f, [n])
        Rmiu(-2:a, -2:b, -2:c) => f
    case (21)
        call c_f_pointer(p, f, [n])
        Dst(-2:a, -2:b, -2:c) => f
    end select
end subroutine

subroutine ref_block_scalars(ppf_in, tiny_in, gmax, gmin, flim2, ki_in, itr_in, kdual, kmodel) &
    bind(C, name="ref_block_scalars")
    use iso_c_binding, only: c_double, c_int
    use Global, only: PPf, Tiny, gama_max, gama_min, F_limiter2, KI, itr_chld, Kind_Dual, Kind_Model
    implicit none
    real(c_double), value :: ppf_in, tiny_in, gmax, gmin, flim2
    integer(c_int), value :: ki_in, itr_in, kdual, kmodel
    PPf = ppf_in
    Tiny = tiny_in
    gama_max = gmax
    gama_min = gmin
    F_limiter2 = flim2
    KI = ki_in
    itr_chld = itr_in
    Kind_Dual = kdual
    Kind_Model = kmodel
end subroutine

! The module variables TIME_RK_V_Update_New reads besides those of ref_block_scalars.
subroutine ref_rk_scalars(a1_in, beta_star_in, prod_lmt_in, ref_in, rxm2_in, csthlnd_in, csth1_in, &
                          if_euler_in, if_turb_in, kind_model_in, kitm_in) bind(C, name="ref_rk_scalars")
    use iso_c_binding, only: c_double, c_int
    use Global, only: A1, Beta_Star, Prod_Lmt, Ref, RXM2, Csthlnd, Csth1, If_Euler, IF_Turb, Kind_Model, KItm
    implicit none
    real(c_double), value :: a1_in, beta_star_in, prod_lmt_in, ref_in, rxm2_in, csthlnd_in, csth1_in
    integer(c_int), value :: if_euler_in, if_turb_in, kind_model_in, kitm_in
    A1 = a1_in
    Beta_Star = beta_star_in
    Prod_Lmt = prod_lmt_in
    Ref = ref_in
    RXM2 = rxm2_in
    Csthlnd = csthlnd_in
    Csth1 = csth1_in
    If_Euler = if_euler_in
    IF_Turb = if_turb_in
    Kind_Model = kind_model_in
    KItm = kitm_in
end subroutine

! The block's residual norms and their locations as TIME_RK and TIME_RK_V_Update_New leave them:
! r = DRSa, DUa, DRSk, DUk, DRSo, DUo; m = Ima, Jma, Kma, Imk, Jmk, Kmk, Imo, Jmo, Kmo.
subroutine ref_block_resid(r, m) bind(C, name="ref_block_resid")
    use iso_c_binding, only: c_double, c_int
    use ref_block_state, only: blk
    implicit none
    real(c_double), intent(out) :: r(6)
    integer(c_int), intent(out) :: m(9)
    r = [blk%DRSa, blk%DUa, blk%DRSk, blk%DUk, blk%DRSo, blk%DUo]
    m = [blk%Ima, blk%Jma, blk%Kma, blk%Imk, blk%Jmk, blk%Kmk, blk%Imo, blk%Jmo, blk%Kmo]
end subroutine

! 1 CONVECT_MUSCL, 2 CONVECT_MUSCL2, 3 CONVECT_MDCD, 4 CONVECT_MDCDHYB, 5 TURBTWO_CONVECT_MUSCL,
! 6 CONVECT_AUSMPW_Plus, 7 TURBTWO_CONVECT_Upwind, 8 TIME_Update_U, 9 TIME_RK_Save, 10 TIME_RK,
! 11 TIME_RK_V_Update_New
subroutine ref_block_call(which) bind(C, name="ref_block_call")
    use iso_c_binding, only: c_int
    implicit none
    integer(c_int), value :: which
    select case (which)
    case (1)
        call CONVECT_MUSCL
    case (2)
        call CONVECT_MUSCL2
    case (3)
        call CONVECT_MDCD
    case (4)
        call CONVECT_MDCDHYB
    case (5)
        call TURBTWO_CONVECT_MUSCL
    case (6)
        call CONVECT_AUSMPW_Plus
    case (7)
        call TURBTWO_CONVECT_Upwind
    case (8)
        call TIME_Update_U
    case (9)
        call TIME_RK_Save
    case (10)
        call TIME_RK
    case (11)
        call TIME_RK_V_Update_New
    end select
end subroutine
