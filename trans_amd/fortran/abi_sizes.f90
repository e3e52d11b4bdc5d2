! Prints the byte sizes of the bind(C) types of trans_hip_mod, one per line, so that a test
! can compare them with the C structs of include/trans_hip.h.
program abi_sizes
    use iso_c_binding
    use trans_hip_mod
    implicit none
    type(trans_params) :: a
    type(trans_patch) :: b
    type(trans_block_desc) :: c
    type(trans_fview) :: d
    type(trans_geom_view) :: e
    type(trans_state_view) :: f
    type(trans_resid) :: g
    type(trans_plane_integrals) :: h
    type(trans_srv_config) :: q1
    type(trans_srv_request) :: q2
    type(trans_srv_status) :: q3
    type(trans_geom_metrics_out) :: q4
    print '(a, i0)', "trans_params ", c_sizeof(a)
    print '(a, i0)', "trans_patch ", c_sizeof(b)
    print '(a, i0)', "trans_block_desc ", c_sizeof(c)
    print '(a, i0)', "trans_fview ", c_sizeof(d)
    print '(a, i0)', "trans_geom_view ", c_sizeof(e)
    print '(a, i0)', "trans_state_view ", c_sizeof(f)
    print '(a, i0)', "trans_resid ", c_sizeof(g)
    print '(a, i0)', "trans_plane_integrals ", c_sizeof(h)
    print '(a, i0)', "trans_srv_config ", c_sizeof(q1)
    print '(a, i0)', "trans_srv_request ", c_sizeof(q2)
    print '(a, i0)', "trans_srv_status ", c_sizeof(q3)
    print '(a, i0)', "trans_geom_metrics_out ", c_sizeof(q4)
end program abi_sizes
