! TEST INFRASTRUCTURE ONLY.  C-callable wrappers over routines of the reference that compile
! unmodified from its own sources: the low-Mach preconditioning (getPrecondPara, PrecondDW;
! Common/UTIL_Precond_Funcs.F90:7-31, 91-129), the MUSCL limiters (ALIMITER, MUSCL2;
! Space/UTIL_MUSCL.F90), the AUSMPW+ face flux (CONVECT_AUSMPW_Plus_Common;
! Space/CONVECT_AUSMPW_Plus.F90:204-291) and the hybrid RANS/LES functions (FSST,
! fd_function, fe_function; Time/Hybrid_function.F90).  oracle/Makefile (target ref) compiles
! those files where they lie, with default REAL promoted to 8 bytes as the FP64 build of the
! reference is, and links them with this driver into oracle/_ref/libtrans_ref.so.  Only tests/
! load that library, to check the oracle's restatements against the reference bit for bit.
!
! The routines read module variables of Module_Global.F90 (PPf, Tiny, XM, Csthlnd, Csth1, Ref,
! Kind_SubHY); each wrapper sets the ones its routines read before the calls.

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
