/*
 * trans_hip.h -- C-ABI of the MI355X (gfx950) residual-and-update hot path of the
 * TRANS block-structured RANS solver.
 *
 * The reference calls parameterless Fortran subroutines on module state, one block at a
 * time, after GLOBAL_SetPointers(iBlock) (main.F90:65-197, Module_Global.F90:274-392).
 * This library replaces that loop body.  A host (the Fortran driver in
 * trans_amd/fortran/ through iso_c_binding, or Python through ctypes) keeps the
 * reference's block arrays in their own layout -- Fortran column-major, component index
 * first, explicit lower bounds (Input/Read_BlockDimensions.F90:306-401) -- and hands
 * them over as trans_fview descriptors.  The library transposes them once into its
 * device layout and runs every stage on the GPU.
 *
 * Every function returns 0 on success, a negative value for an API/HIP error and a
 * positive value for a numerical failure; trans_hip_last_error() gives the text.
 * A context is not thread safe; all device work of a context runs on one stream.
 * No pointer passed in is retained after the call returns.
 */
#ifndef TRANS_HIP_H
#define TRANS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRANS_ML 7 /* rho, rho*u, rho*v, rho*w, rho*E, rho*k, rho*omega (Module_Global.F90:13) */

/* Error codes */
#define TRANS_OK 0
#define TRANS_EINVAL (-1)
#define TRANS_EHIP (-2)
#define TRANS_ENOMEM (-3)
#define TRANS_EUNSUPPORTED (-4)
#define TRANS_ESCRATCH (-5) /* input larger than a static staging array (trans_geom_wall_distance_static) */
#define TRANS_ETIMEOUT (-6) /* a bounded wait of the tile-wavefront sweep expired (results invalid) */
#define TRANS_ENONFINITE 1 /* a residual norm came out NaN (Time/TIME_Update_V.F90:124-151) */

/*
 * Deck scalars the hot path reads from module Global.  Filled by the host from
 * original.org / dimension_global.dat exactly as INPUT_ReadOrg (Input/ReadOrg.F90:98-160),
 * ParameterSet and NonDimensionlization (Input/ParameterSet.F90:1-96) do.
 */
typedef struct trans_params {
    /* switches */
    int kind_spatial;    /* 1 = MUSCL2, 2 = MUSCL/ALIMITER, 3 = MDCD, 4 = MDCDHYB */
    int kind_roe;        /* 1 = Roe, 2 = rotated Roe, 3 = AUSMPW+ */
    int if_euler;
    int if_turb_org;
    int kind_model;      /* 1 = SST, 2 = WD+ k-omega */
    int kind_dual;       /* 0, 1, 2 */
    int kchld;           /* sub-iterations per outer iteration */
    int if_rk;           /* 0 = LU-SGS, 1 = SSP-RK3 */
    int kitm;            /* first iteration with turbulence on */
    int kind_hybrid;     /* 0 = RANS, 3 = DES/DDES (any value > 0 keeps k-omega on from KI = 1) */
    int kind_subhy;      /* with Kind_Hybrid 3: 0 DES, 1/3 DDES with Menter's F1, 2 DDES with F2, 5 IDDES */
    int if_comppress;    /* SST compressibility correction */
    int kind_cc;
    int kind_rotmod;     /* 0 or 2 (mode 2 only writes f_r2, which nothing reads; mode 1 reads an unset index) */
    int if_precondition; /* 0, or 1: low-Mach preconditioning (hyperplane LU-SGS sweep); fixed at creation */
    int cfl_sub_step;    /* >= 1 (SpectralRadius.F90:159 divides by it) */
    int ki_ini;
    int implicit_fp32;   /* 1: LU-SGS operator and Rds diagonals in FP32 */
    /* nondimensional constants */
    double xm, rxm2, ppf;      /* M, gamma*M^2, 1/(gamma*M^2) */
    double csthlnd, csth1;     /* Sutherland: 110/Tinf, Tinf/273.15 */
    double ref;                /* Reynolds number per unit grid length */
    double omega1;             /* rotation rate about x (nondimensional) */
    double tiny;               /* 1e-20 (main.F90:12) */
    double vxf, vyf, vzf;      /* freestream direction */
    double dt_lusgs_main, cfl_sub_start, cfl_sub_end;
    double gama_max, gama_min; /* MDCD dissipation bounds */
    double prod_lmt, visratio, fsti;
    double a1, akapa, beta_star;
    double sigk1, sigo1, beta1, gama1;
    double sigk2, sigo2, beta2, gama2;
    double p0in, t0in, pouthub, ps00, ts00;
    double rkf, rof, rmiuf;
    double f_limiter1, f_limiter2; /* deck: DES blending floor; MDCDHYB switch to MUSCL at sensor >= f_limiter2 */
    double alscale;                /* deck: DES length-scale floor of the blending function */
} trans_params;

/*
 * One BC patch: the fields of ConnectivityStruct (Module_Block.F90:4-23) the hot path
 * reads, as PLOT3D_Read_And_Init builds them (Input/PLOT3DReadInit.F90:135-348).
 * Indices keep the reference's 1-based point/cell numbering; target_block and
 * target_conn are 0-based positions in the arrays given to trans_hip_create.
 */
typedef struct trans_patch {
    int bctype;          /* 1 wall, 2 far field, 3 symmetry, 5 inlet, 7 outlet, 9 interface */
    int subbctype;
    int pertype;         /* periodic multiple (0 = plain interface) */
    int point_start[3];
    int point_end[3];
    int ijkdirect;       /* 1, 2, 3 */
    int left_or_right;   /* 1 = min face, 2 = max face */
    int ijkadd[3];
    int cijkadd[3], cijkadd1[3], cijkadd2[3];
    int ceijkadd[3], ceijkadd1[3], ceijkadd2[3];
    int target_block;
    int target_conn;
    int transform[3];
    int transform_t[3];
    int rev_matrix[3];
    int lcross;
    double rotate_angle; /* rotateAngle(1) */
} trans_patch;

typedef struct trans_block_desc {
    int ni, nj, nk; /* points; cells are ni-1, nj-1, nk-1 */
    int npatch;
    const trans_patch* patches;
} trans_block_desc;

/*
 * A Fortran array A(ncomp, lo0:hi0, lo1:hi1, lo2:hi2) in column-major order.
 * Element (c, i, j, k), c 0-based, is at
 *   p[c + ncomp*((i-lo0) + n0*((j-lo1) + n1*(k-lo2)))],  n0 = hi0-lo0+1, n1 = hi1-lo1+1.
 * For SD(3,3,...) ncomp = 9 and c = (dir-1) + 3*(xyz-1).
 */
typedef struct trans_fview {
    double* p;
    int ncomp;
    int lo[3];
    int hi[3];
} trans_fview;

/* Geometry of one block (Geom/CalcDst.F90:28-719, Geom/startup.F90:396-755). */
typedef struct trans_geom_view {
    trans_fview sd;    /* SD(3,3,...) face area vectors */
    trans_fview grad;  /* Grad(3,...) face areas */
    trans_fview alagm; /* Alagm(3,...) face-to-face lengths */
    trans_fview vol;   /* Vol */
    trans_fview rad;   /* rad = sqrt(yc^2+zc^2) */
    trans_fview dst;   /* wall distance */
    trans_fview aleng; /* largest neighbour distance (DES length) */
    trans_fview aseng; /* smallest neighbour distance (IDDES; may be unset otherwise) */
    trans_fview yc, zc;
    trans_fview thtc;  /* circumferential angle atan2(zc, yc) (outlet radial equilibrium) */
    /* MARKWALLi0(NJ1,NK1), MARKWALLim, j0(NI1,NK1), jm, k0(NI1,NJ1), km; NULL = none */
    const int* markwall[6];
} trans_geom_view;

/* Flow state of one block. wt2 may have p == NULL. */
typedef struct trans_state_view {
    trans_fview v;    /* V(7,...) conservative variables */
    trans_fview pp;   /* PP */
    trans_fview t;    /* T */
    trans_fview rmiu; /* eddy viscosity */
    trans_fview wt0;  /* dual-time history (7,...) */
    trans_fview wt1;
    trans_fview wt2;
} trans_state_view;

/* Per-block norms stashed by TIME_Update_V (Time/TIME_Update_V.F90:892-908). */
typedef struct trans_resid {
    double drsa, dua;
    int ima, jma, kma;
    double drsk, duk;
    int imk, jmk, kmk;
    double drso, duo;
    int imo, jmo, kmo;
} trans_resid;

/* Fields readable through trans_hip_get_field (parity hooks). */
#define TRANS_FIELD_D 0      /* D(7,...) residual, then the LU-SGS increment */
#define TRANS_FIELD_RDS 1    /* Rds(3,...) */
#define TRANS_FIELD_DTM 2    /* Dtm */
#define TRANS_FIELD_DQDXYZ 3 /* DQdxyz(18,...) */
#define TRANS_FIELD_SHOCK 4  /* shock sensor */
#define TRANS_FIELD_FSCHEME 5 /* DES blending Fscheme of the rotated Roe flux (Kind_Roe = 2) */

/* Integrals over one constant-index face plane of a block (all faces of direction dir at
 * face index `index`, i.e. between cells index-1 and index).  The face state is the
 * reference's mass-monitor weighting of the two adjacent cells
 * (Output/OUTPUT_Residual_NS.F90:19-32): q_f = (Vol(index-1) q(index) + Vol(index) q(index-1))
 * / (Vol(index-1) + Vol(index)); dm = (rho u)_f . S.  Total pressure and temperature use the
 * reference's output formula p0 = p (1 + 0.2 M^2)^3.5 (Output/Average_Field.F90:66) and
 * T0 = T (1 + 0.2 M^2), with the stored (relative-frame) velocity.  The absolute-frame
 * averages (no counterpart in the reference; the rotor's stage performance needs them) use
 * the absolute velocity u, v - omega1 z_f, w + omega1 y_f, where y_f, z_f are the same
 * weighting of the cell centres; they equal the relative ones when omega1 = 0.
 * Nondimensional, as the state is. */
typedef struct trans_plane_integrals {
    double mass_flow;   /* sum of dm (the reference's msum before its Roref*Vref scaling) */
    double area;        /* sum of |S| */
    double p0_mass_avg; /* sum(p0 dm) / sum(dm), relative frame */
    double t0_mass_avg; /* sum(T0 dm) / sum(dm), relative frame */
    double p0_abs_mass_avg; /* the same with the absolute velocity */
    double t0_abs_mass_avg;
} trans_plane_integrals;

typedef struct trans_ctx trans_ctx;

/* Allocate device state for all blocks.  Replaces INPUT_Read_BlockDimensions'
 * ALLOCATEs (Input/Read_BlockDimensions.F90:306-401) and PLOT3D_Read_And_Init's
 * connectivity (Input/PLOT3DReadInit.F90:1-349), which the host has already built.
 * TRANS_EINVAL for a block below the stencils (ni, nj < 4 or nk < 3 points) or beyond the
 * kernels' 32-bit per-block indexing (more than 2^28 padded cells). */
int trans_hip_create(const trans_params* params, int nblk, const trans_block_desc* blocks,
                     trans_ctx** out);
/* (The LU-SGS sweeps run as a line march when every block's cross-section fits one workgroup,
 * else as a hyperplane wavefront; TRANS_HIP_LINE_SWEEP=0 in the environment at create time
 * selects the hyperplane wavefront for all blocks.  Both give the same results.) */
int trans_hip_destroy(trans_ctx* ctx);
const char* trans_hip_last_error(const trans_ctx* ctx);

/* Replace the deck scalars (original.org is re-read every 10 iterations, main.F90:49-52). */
int trans_hip_set_params(trans_ctx* ctx, const trans_params* params);
/* KI, the outer iteration counter (Module_Global.F90:79). */
int trans_hip_set_iteration(trans_ctx* ctx, int ki);

/* Upload geometry (after GEOM_CalcDSTs / GeomTransform / GeomTransform1, main.F90:28-31). */
int trans_hip_set_geometry(trans_ctx* ctx, int blk, const trans_geom_view* g);
/* Upload / download the flow state including all ghost layers the views cover. */
int trans_hip_set_state(trans_ctx* ctx, int blk, const trans_state_view* s);
int trans_hip_get_state(trans_ctx* ctx, int blk, const trans_state_view* s);
int trans_hip_get_field(trans_ctx* ctx, int blk, int which, const trans_fview* out);

/* One call per reference routine, over all blocks of the context: */
int trans_hip_spectral_radius(trans_ctx* ctx);     /* TIME_SpectralRadius, main.F90:66-69 */
int trans_hip_bc(trans_ctx* ctx, int sweep);       /* BC_Boundarys(sweep), main.F90:72,84 */
int trans_hip_space(trans_ctx* ctx);               /* main.F90:74-133 except the two BC calls */
int trans_hip_residual(trans_ctx* ctx);            /* main.F90:66-133: SpectralRadius..convection */
int trans_hip_lusgs_update(trans_ctx* ctx, trans_resid* per_block);
                                                   /* TIME_LUSGS + TURBTWO_TIME_LUSGS_SST +
                                                      TIME_Update_V, main.F90:137-151 */
/* Stage itr_chld (1, 2 or 3) of the Runge-Kutta update (IF_RK = 1): TIME_RK_Save + TIME_RK +
 * TURBTWO_TIME_LUSGS_SST + TIME_RK_V_Update_New, main.F90:158-170.  per_block (may be NULL)
 * receives DRSa from the density residual before the stage and DRSk/DRSo from the k-omega
 * increments. */
int trans_hip_rk_update(trans_ctx* ctx, int itr_chld, trans_resid* per_block);
int trans_hip_update_u(trans_ctx* ctx);            /* TIME_Update_U, main.F90:192-197 */
/* n outer iterations of main.F90:48-211 (Kchld sub-iterations, or the three Runge-Kutta stages
 * when IF_RK = 1; update of Wt0/Wt1, KI += 1).
 * per_block (nblk entries, may be NULL) receives the norms of the last sub-iteration. */
int trans_hip_iterate(trans_ctx* ctx, int n_iter, trans_resid* per_block);
/* Plane integrals (struct above) of block blk (0-based), direction dir (1 i, 2 j, 3 k), face
 * index 1..N_dir.  Generalises OUTPUT_Residual_NS's mass monitor, which the reference
 * hard-wires to i = 4 of block 1 and i = NI-4 of block 2 (Output/OUTPUT_Residual_NS.F90:15-60);
 * the ratio and efficiency of two planes follow on the host. */
int trans_hip_plane_integrals(trans_ctx* ctx, int blk, int dir, int index, trans_plane_integrals* out);
/* Wait for all queued device work of the context. */
int trans_hip_sync(trans_ctx* ctx);

/* Kernel timing (no counterpart in the reference, which has no timers): when on, every
 * kernel launch of the context is bracketed by HIP events on the context's stream.
 * trans_hip_profile(ctx, 1) also clears the totals.  profile_get returns the summed
 * device time and the launch count of the idx-th kernel name. */
int trans_hip_profile(trans_ctx* ctx, int on);
int trans_hip_profile_count(trans_ctx* ctx);
int trans_hip_profile_get(trans_ctx* ctx, int idx, char* name, int name_len, double* total_ms,
                          long* launches);

/* ---- Restart records and time statistics ------------------------------------------------
 * Files are Fortran unformatted sequential, as the reference writes them: every record framed
 * by its byte count as a 4-byte integer before and after (records below 2 GiB). */

/* OUTPUT_SaveFields for block blk (Output/OUTPUT_SaveField.F90:14-26): records KI (int32),
 * V(1:7), Rmiu, then Wt0 (Kind_dual 1) or Wt0 and Wt1 (Kind_dual 2), each over the interior
 * cells with K fastest, then J, I and the component outermost, as float64.  ki is the KI the
 * record carries (the reference writes the iteration just completed). */
int trans_hip_write_restart(trans_ctx* ctx, int blk, int ki, const char* path);
/* Reads such a record into block blk and recomputes p and T from V (Input/Init.F90:22-46,
 * 146-160); *ki (may be NULL) receives the stored KI (the reference continues at KI + 1).
 * The Kind_Hybrid history reset of Input/Init.F90:54-71 is the host's. */
int trans_hip_read_restart(trans_ctx* ctx, int blk, const char* path, int* ki);

/* Time statistics of Average_Fields (Output/Average_Field.F90:46-263), TRANS_NSTAT
 * components per cell in the order of its FLfld record: U V W R M Mu k P0 P T uu vv ww uv uw vw
 * (running means from KI_aver on; the last six are the modelled Reynolds stresses) and
 * uurms vvrms wwrms uvrms vwrms uwrms pprms (from KI_statis on).  stats_enable allocates
 * them (zero) and makes trans_hip_iterate update them after every outer iteration with
 * KI >= ki_aver, as main.F90:185-188 does; ki_aver < 0 switches that off.
 * trans_hip_average_fields is the routine-level call for hosts that drive the loop
 * routine by routine (not to be combined with trans_hip_iterate's own update). */
#define TRANS_NSTAT 23
int trans_hip_stats_enable(trans_ctx* ctx, int ki_aver, int ki_statis);
int trans_hip_average_fields(trans_ctx* ctx);
int trans_hip_get_stats(trans_ctx* ctx, int blk, const trans_fview* out);
int trans_hip_set_stats(trans_ctx* ctx, int blk, const trans_fview* in);
/* FLfld record file (Output/Average_Field.F90:264-281; read back at Input/Init.F90:77-91): one
 * record per interior cell, I fastest, of Xc Yc Zc (from the caller's cell-centre views) and
 * the TRANS_NSTAT statistics. */
int trans_hip_write_stats(trans_ctx* ctx, int blk, const trans_fview* xc, const trans_fview* yc,
                          const trans_fview* zc, const char* path);
int trans_hip_read_stats(trans_ctx* ctx, int blk, const char* path);

/* ---- Tecplot ASCII field output ------------------------------------------------------------
 * Output/OUTPUT_InstField.F90:150-370 (kind_field = 0) writes each block's point coordinates
 * and 31 cell-centred variables through the closed TecIO library; this writes the same
 * variables in Tecplot's documented ASCII format: TITLE, VARIABLES = "X" "Y" "Z" density U V
 * W Uabs Vabs Wabs Ma Mabs Ps P0rel P0 Ts T0rel T0 Ox Oy Oz Om Q Yita4 Rmiu k shock fr uu vv
 * ww uv uw vw, one ordered zone I=NI, J=NJ, K=NK with DATAPACKING=BLOCK and the 31 field
 * variables CELLCENTERED, values in i-fastest order, five per line as %.10e.
 * x, y, z: the block's point coordinates XX, YY, ZZ (views covering 1..NI, 1..NJ, 1..NK);
 * scales: {Roref, Vref, Pref, Tinf} as the reference multiplies them in (NULL: nondimensional).
 * The variables follow the reference's formulas, including its absolute-frame quirks (Mabs is
 * |c| unscaled, P0 = Ps (1 + Mabs^2) * 3.5, T0 = Ts (1 + Mabs^2)); fr is 1 (no rotation-
 * curvature model); velocity gradients are the DQdxyz of the last residual evaluation. */
int trans_hip_write_tecplot(trans_ctx* ctx, int blk, const trans_fview* x, const trans_fview* y,
                            const trans_fview* z, const double* scales, const char* path);
/* The host-side formatter of the same file, usable without a GPU: point arrays x, y, z of
 * ni*nj*nk values and ncellvar cell-centred arrays of (ni-1)(nj-1)(nk-1) values (cellvars,
 * variable-major), all i fastest; names may not contain '"'. */
int trans_tecplot_write_ascii(const char* path, const char* title, const char* zone, int ni, int nj, int nk,
                              const double* x, const double* y, const double* z, int ncellvar,
                              const char* const* cellvar_names, const double* cellvars);

/* ---- Launch autotuner (no counterpart in the reference) ------------------------------------
 * Times `iters` iterations of every launch configuration of the context from the same saved
 * state and keeps the fastest: the LU-SGS sweep (sweep = 0, 1, 2: line march along i, j or k,
 * where the cross sections fit one workgroup; 3: tile wavefront over many CUs with the tile
 * shapes tile_a x tile_b and progress chunks below; -2: one launch per hyperplane over all
 * blocks; -1: one workgroup per block walking its hyperplanes), helper streams for the
 * face kernels (streams = 4) or one stream (1), HIP-graph replay (graph = 1) or eager launches.
 * Every configuration computes the same iteration (the sweep variants contract multiply-adds
 * in different places, so they agree within rounding, not bitwise).  The state and KI are restored;
 * out (max_out entries, may be NULL) receives the timings, *n_out their count.  The sweep
 * buffer is reallocated for the chosen axis, so the context's device bytes may then differ
 * from trans_hip_footprint's.  get_tuning reports the configuration in use. */
typedef struct trans_tune_entry {
    int sweep;     /* 0, 1, 2 line march along i, j, k; 3 tile wavefront; -2 one launch per hyperplane;
                      -1 one workgroup per block; -3 plain drop-in schedule (set_tuning only) */
    int streams;
    int graph;
    double ms_per_iter;
    int tile_a, tile_b; /* sweep 3: lines of a tile along i and along k (tile_a * tile_b <= 256) */
    int chunk;          /* sweep 3: steps between a tile's progress stores */
} trans_tune_entry;
int trans_hip_autotune(trans_ctx* ctx, int iters, trans_tune_entry* out, int max_out, int* n_out);
int trans_hip_get_tuning(const trans_ctx* ctx, trans_tune_entry* out);
/* Applies a configuration (ms_per_iter ignored), e.g. one found by an earlier autotune.
 * sweep = -3 selects the plain drop-in schedule (streams, graph and the tile fields ignored):
 * every routine of trans_hip_iterate is launched once per block, as a straight port of the
 * reference's per-block calls would run, each LU-SGS sweep one launch per hyperplane per
 * block (with IF_RK = 1: the stage kernels once per block, the k-omega sweeps one launch per
 * hyperplane per block), on one stream, without graph replay; the same arithmetic as the
 * batched schedule.  Decks without preconditioning only (TRANS_EUNSUPPORTED otherwise).  A
 * baseline for the batched schedule; autotune leaves it for the batched candidates.
 * A tile-wavefront sweep whose progress wait exceeds its spin limit (a schedule that cannot
 * advance) ends and makes the next synchronising call return TRANS_ETIMEOUT. */
int trans_hip_set_tuning(trans_ctx* ctx, const trans_tune_entry* t);

/* ---- Geometry pre-processing on the device -------------------------------------------------
 * Wall distance of Geom/CalcDst.F90:121-158 (IF_dst = 0): for each of ncell cell centres
 * (cells: x[ncell], then y, then z) the minimum over npanel wall-panel centres (panels: same
 * layout) of the squared distance, dst = sqrt(max(min, 1e-20)); 1e3 everywhere without
 * panels.  Host pointers; no context needed.  Bitwise the same as the host computation. */
int trans_geom_wall_distance(const double* cells, long ncell, const double* panels, long npanel, double* dst);

/* trans_geom_wall_distance without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, serving loops
 * that must not call hipMalloc, harnesses that forbid it).  Same arguments and bitwise the same
 * result as trans_geom_wall_distance, but cells, panels and the squared minima are staged in a
 * static device array of TRANS_WD_STATIC_DOUBLES doubles (32 MiB; one copy per GPU, held until
 * the process ends), and the work runs on the null stream of the current device.  Requires
 * 4 ncell + 3 npanel <= TRANS_WD_STATIC_DOUBLES, else returns TRANS_ESCRATCH without touching
 * the device.  Calls that overlap on one GPU (from several host threads) share the array and
 * race; serialise them. */
#define TRANS_WD_STATIC_DOUBLES (4L << 20)
int trans_geom_wall_distance_static(const double* cells, long ncell, const double* panels, long npanel, double* dst);

/* Outputs of trans_geom_metrics: host arrays of one block in the padded layout (bounds
 * -2:N+2 on each axis, i fastest, component first for sd (9: direction fastest, then x, y,
 * z), grad (3) and alagm (3)). */
typedef struct trans_geom_metrics_out {
    double *xc, *yc, *zc; /* cell centres; boundary face centres at index 0 and N */
    double *thtc, *rad;   /* atan2(zc, yc), sqrt(yc^2 + zc^2) */
    double *aleng, *aseng;/* largest / smallest distance to the six neighbour centres */
    double *vol;
    double *sd, *grad, *alagm;
} trans_geom_metrics_out;

/* The O(N) metrics of Geom/CalcDst.F90:170-632 for one block of ni x nj x nk points (xx, yy,
 * zz in the padded layout, points at indices 1..N): centres, face area vectors and areas
 * (plus tiny under the root), face-to-face lengths Alagm, neighbour distances, volumes, and
 * the first-ghost copies of Vol, Alagm, Grad and the SD extension to faces 0 and N+1.
 * Entries the routine does not set keep the caller's values; the wall distance is
 * trans_geom_wall_distance.  Bitwise the same as the reference's arithmetic order (thtc up
 * to the device atan2).  TRANS_EINVAL on a cell volume below tiny (as CalcDst stops).
 * Host pointers; no context needed. */
int trans_geom_metrics(int ni, int nj, int nk, const double* xx, const double* yy, const double* zz, double tiny,
                       const trans_geom_metrics_out* out);

/* Device bytes a context for these inputs would allocate (trans_hip_create's own sizing,
 * run without allocating or touching the device), and the bytes a context holds.  No
 * counterpart in the reference; used to admit solves by HBM footprint. */
int trans_hip_footprint(const trans_params* params, int nblk, const trans_block_desc* blocks,
                        size_t* bytes);
int trans_hip_device_bytes(const trans_ctx* ctx, size_t* bytes);

/* ---- Serving: continuous batching of independent steady solves -------------------------
 * No counterpart in the reference, which runs one case per process.  A server admits
 * submitted solves, in arrival order and first fit, while their footprints fit hbm_budget
 * and fewer than max_active are resident; each resident solve keeps its state in HBM and
 * advances in chunks of `chunk` iterations on its own stream, so the chunks of all resident
 * solves overlap on the device.  A solve finishes when its global residual
 * DRSa = sqrt(sum_b N_b DRSa_b^2 / sum_b N_b) (N_b cells of block b) has dropped to
 * target_drop times its first-iteration value, or after max_iter iterations; it then leaves
 * the device and its final state waits on the host for trans_srv_result.  The iteration that
 * first reached the target is reported exactly; the state is the one at the end of that
 * chunk.  A solve's iterations depend only on its own inputs and on `chunk`, so its result
 * is the same alone and in any mix.  One host thread per server. */
typedef struct trans_srv trans_srv;

typedef struct trans_srv_config {
    size_t hbm_budget; /* device bytes the resident solves may hold together */
    int max_active;    /* most solves resident at once */
    int chunk;         /* iterations queued per solve between convergence checks */
} trans_srv_config;

typedef struct trans_srv_request {
    trans_params params;
    int nblk;
    const trans_block_desc* blocks; /* nblk */
    const trans_geom_view* geom;    /* nblk */
    const trans_state_view* state;  /* nblk: initial state */
    double target_drop;             /* e.g. 1e-8 */
    int max_iter;
} trans_srv_request;

#define TRANS_SRV_QUEUED 0
#define TRANS_SRV_RUNNING 1
#define TRANS_SRV_CONVERGED 2
#define TRANS_SRV_MAX_ITER 3
#define TRANS_SRV_CANCELLED 4
#define TRANS_SRV_FAILED 5

typedef struct trans_srv_status {
    int state;                /* TRANS_SRV_* */
    int error;                /* TRANS_E* when FAILED */
    int iterations;           /* iterations run */
    int iterations_to_target; /* first iteration at or below the target drop, 0 = not reached */
    double drsa_first;        /* global DRSa of iteration 1 */
    double drsa_last;         /* global DRSa of the last iteration run */
    double t_submit, t_admit, t_done; /* seconds since trans_srv_create; 0 = not yet */
    size_t footprint;         /* device bytes while resident */
} trans_srv_status;

int trans_srv_create(const trans_srv_config* cfg, trans_srv** srv);
/* Copies the request's inputs (the caller's arrays are not used after the call); rejects a
 * request the solver rejects or that could never fit the budget. */
int trans_srv_submit(trans_srv* srv, const trans_srv_request* req, int* id);
/* One scheduling round: collect finished chunks, retire finished solves, admit waiting ones,
 * queue the next chunks.  wait != 0 first blocks until the oldest queued chunk has ended.
 * n_unfinished (may be NULL) receives the number of queued or running solves. */
int trans_srv_progress(trans_srv* srv, int wait, int* n_unfinished);
int trans_srv_poll(trans_srv* srv, int id, trans_srv_status* out);
int trans_srv_cancel(trans_srv* srv, int id);
/* Final state of block blk (0-based) of a finished solve into the caller's views. */
int trans_srv_result(trans_srv* srv, int id, int blk, const trans_state_view* out);
/* Drops the host copy of a finished solve's state. */
int trans_srv_release(trans_srv* srv, int id);
int trans_srv_destroy(trans_srv* srv);
const char* trans_srv_last_error(const trans_srv* srv);

#ifdef __cplusplus
}
#endif
#endif /* TRANS_HIP_H */
