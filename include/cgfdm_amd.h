/* C ABI of the MI355X CGFDM3D-EQR-SC time-stepping library (libcgfdm_amd.so).
 *
 * The reference has no library interface: its boundary for this path is the set
 * of host launchers that propagate() calls once per step and per RK stage
 * (reference src/propagate.cpp:186-321, declared in src/functions.h:88-216),
 * plus the setup stage that builds their inputs (src/run.cpp:56-117). Each entry
 * point below names the reference functions it replaces. Differences by design:
 *   - an opaque handle instead of globals (reference run.cpp:20-24);
 *   - int status codes instead of MPI_Abort (same numeric codes);
 *   - one GPU, one rank: the domain is the whole grid at its real shape, so
 *     the halo exchange (MPI_send_recv_FLOAT.cpp:388-505) disappears;
 *   - work is enqueued on one HIP stream (cgfdm_set_stream); a step may fork
 *     its surface-layer kernels onto an internal stream, joined back to that
 *     stream with events before the step's later work, so everything a call
 *     enqueues is complete once the solver stream is; calls are not thread
 *     safe on the same handle.
 * Arrays are host memory unless stated otherwise. Grid arrays are SoA over the
 * padded grid, x fastest: a[c][k][j][i], NX = nx + 6 etc. (HALO = 3).
 */
#ifndef CGFDM_AMD_H
#define CGFDM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (reference MPI_Abort codes where one exists) */
#define CGFDM_OK 0
#define CGFDM_ERR_ARG 2          /* invalid argument or call order (no reference counterpart) */
#define CGFDM_ERR_HIP 3          /* HIP runtime error (reference CHECK, macro.h:94-103, prints and continues) */
#define CGFDM_ERR_NO_PARAMS 100  /* params.json missing (getParams.cpp:25-29) */
#define CGFDM_ERR_PARSE 1001     /* params.json / station.json parse error (getParams.cpp:48-53) */
#define CGFDM_ERR_CFL 110        /* DT > 1.34 Hmin / Vp_max (calc_CFL.cpp:325,341-345) */
#define CGFDM_ERR_PML 130        /* nPML >= local n (pml_deriv.cpp:55-85) */
#define CGFDM_ERR_SOURCE 180     /* inconsistent source description (modelChecking.cpp:23-50) */
#define CGFDM_ERR_ALLOC 10001    /* device allocation failure (wave_deriv.cpp:49-53) */
#define CGFDM_ERR_DISABLED 4     /* code path switched off in this build (no reference counterpart):
                                    CGFDM stepping, forward_record and adjoint */

#define CGFDM_SCHEME_CGFDM 0     /* Makefile_CGFDM: MacCormack + RK4 + CFS-PML + traction image */
#define CGFDM_SCHEME_AWENO 1     /* Makefile_SCFDM: WENO/LF alternative flux + TVD-RK3 + sponge */

typedef struct cgfdm_solver cgfdm_solver;

typedef struct cgfdm_desc {
  int nx, ny, nz;     /* interior cells of the domain (grid.nx/ny/nz, init_grid.cpp:37-45 with P = 1) */
  int npml;           /* params.nPML (getParams.cpp:137-138); 0 = no absorbing layer. CGFDM only:
                         the AWENO scheme absorbs with the 20-layer exponential sponge of the
                         SCFDM build (exp_decay_absorb.cpp:108-145) and ignores npml */
  int free_surface;   /* 1 = FREE_SURFACE build (traction image on the top face) */
  int scheme;         /* CGFDM_SCHEME_* */
  int precision;      /* 4 = FP32 wavefield (reference FLOAT), 8 = FP64 wavefield */
  int coef_storage;   /* 0 = wavefield precision, 1 = fp16, 2 = bf16 storage of metric and medium
                         planes, with stress carried as T / (rho_ref Vp_ref) (macro.h:75-80 Cs);
                         CGFDM with the FP32 wavefield only */
  int nt;             /* steps the station buffers hold, NT = TMAX / DT (propagate.cpp:88) */
  double dh, dt;      /* params.DH, params.DT */
} cgfdm_desc;

/* Setup: FP64 metrics/Jacobian (contravariant.cpp:182-285; SCFDM: central-difference
 * covariant basis and cofactors, contravariant.cpp:81-177), metric mirror and
 * traction matrices (contravariant.cpp:321-519), (Vs,Vp,rho) -> (mu,lambda,1/rho)
 * and CJM packing (CJM.cpp:53-181), CFL check (calc_CFL.cpp:255-349), PML
 * profiles (init_pml_para.cpp:97-285) and wave/aux allocation (wave_deriv.cpp:31-69,
 * pml_deriv.cpp:17-124). Replaces run.cpp:75-111.
 * coord: float[3][NZ][NY][NX] (x, y, z in metres, as constructCoord, coord.cpp:259-333)
 * medium: float[3][NZ][NY][NX] (Vs, Vp, rho, as constructMedium, medium.cpp:251-317) */
int cgfdm_create(const cgfdm_desc* desc, const float* coord, const float* medium, cgfdm_solver** out);

/* The reference's setup from its input files (run.cpp:28-138 before the time loop):
 * params.json (getParams.cpp:17-242, NULL = "params.json"), station.json
 * (station.cpp:92-301, NULL = "station.json"), the grid (flat, Gaussian hill or the
 * SRTM terrain-following grid, coord.cpp, terrain.cpp:206-860), the medium
 * (homogeneous, layered, Shen or CRUST1.0: medium.cpp, dealMedium.cpp,
 * crustMedium.cpp), the source (finite-fault file, multiSource.cpp:780-868, or a
 * smoothed single source, singleSource.cpp), the stations, and the setup files
 * the reference writes (model slices data_io.cpp:589-618, terrain and lon/lat
 * terrain.cpp:28-118, source files multiSource.cpp:594-672) into params.OUT.
 * nsteps >= 0 replaces NT = TMAX / DT. Errors: 100 no params.json, 1001 parse,
 * 110 CFL, 130 PML, 180 source. */
int cgfdm_create_from_params(const char* params_path, const char* stations_path, int scheme, int precision,
                             int coef_storage, int nsteps, cgfdm_solver** out);

/* Output files of a solver made by cgfdm_create_from_params, in the reference's
 * byte layouts and rank split: CGFDM_OUT_SLICES the Vx, Vy, Vz slices of the
 * current state named by the last step taken, plus the free-surface velocity and
 * displacement slices when sliceFreeSurf is set (data_io.cpp:485-633);
 * CGFDM_OUT_STATIONS station_mpi_X_Y_Z.bin (station.cpp:422-443); CGFDM_OUT_PGV
 * PGV_Z_mpi_X_Y_Z.bin (PGV.cpp:177-185). */
#define CGFDM_OUT_SLICES 1
#define CGFDM_OUT_STATIONS 2
#define CGFDM_OUT_PGV 4
int cgfdm_write_outputs(cgfdm_solver* s, int what);
/* Step count (NT, or the nsteps given) and IT_SKIP of a solver made by
 * cgfdm_create_from_params (propagate.cpp:88, the slice cadence). */
int cgfdm_run_info(cgfdm_solver* s, int* nt, int* it_skip);

/* Finite-fault moment-rate table after solveMomentRate (multiSource.cpp:450-535):
 * rate[npts][nt][6] = (Mxx,Myy,Mzz,Mxy,Mxz,Myz) * area * slip-rate, cell[npts] = padded
 * linear cell index. Points sharing a cell are merged (multiSource.cpp:537-592), the
 * table is made time-major (multiSource.cpp:706-752) and scaled by mu/(J DH^3)
 * (addMoment.cpp:27-98). Replaces init_MultiSource (multiSource.cpp:780-868) after
 * the projection step and calculateMomentRate. */
int cgfdm_set_source(cgfdm_solver* s, int npts, int nt, double dt, const long long* cell, const float* rate);

/* Station cells as padded (i, j, k) triples (station.cpp:92-301 after parsing). */
int cgfdm_set_stations(cgfdm_solver* s, int ns, const int* ijk);

/* Advance nsteps full time steps: source injection at t = it*DT, RK stages with
 * derivative, free surface, PML and update, then station/PGV/displacement
 * recording (propagate.cpp:186-321). AWENO: source as strain rate, three TVD-RK3
 * stages of WENO5/Lax-Friedrichs alternative-flux differences
 * (alternative_flux_FD.cpp:1437-1569, wave_rk_tvd.cpp), characteristic free
 * surface, exponential sponge, observers. Asynchronous on the solver stream. */
int cgfdm_step(cgfdm_solver* s, int nsteps);

/* Use an existing HIP stream (hipStream_t passed as void*); NULL = default stream. */
int cgfdm_set_stream(cgfdm_solver* s, void* stream);
int cgfdm_synchronize(cgfdm_solver* s);
/* number of steps taken so far */
int cgfdm_step_index(cgfdm_solver* s);

/* Station record float[ns][nt][9] (Vx,Vy,Vz,Txx,Tyy,Tzz,Txy,Txz,Tyz), the device
 * buffer that station.cpp:409-443 copies back and writes. */
int cgfdm_get_stations(cgfdm_solver* s, float* out);
/* PGV map float[ny][nx][2] (PGVh, PGV) of the surface layer (PGV.cpp:62-239). */
int cgfdm_get_pgv(cgfdm_solver* s, float* out);
/* Surface displacement float[3][ny][nx] (FreeSurfU*, solve_displacement.cpp:28-74). */
int cgfdm_get_surface_displacement(cgfdm_solver* s, float* out);
/* One wavefield component of the current state, double[NZ][NY][NX] in physical
 * units; source of the slice output (data_io.cpp:485-587). CGFDM: (Vx,Vy,Vz,Txx,
 * Tyy,Tzz,Txy,Txz,Tyz); AWENO: the conservative state (rho vx, rho vy, rho vz,
 * exx, eyy, ezz, gyz, gxz, gxy) (station records and PGV are physical for both). */
int cgfdm_get_field(cgfdm_solver* s, int comp, double* out);
/* One wavefield component of the current state on an interior plane, float in
 * physical units (Vx,Vy,Vz,Txx,Tyy,Tzz,Txy,Txz,Tyz for both schemes; AWENO converts
 * its conservative state as wave_conserv2phy, data_io.cpp:435-483). axis 0: the y-z
 * plane at padded i = index, float[nz][ny]; axis 1: the x-z plane at padded j,
 * float[nz][nx]; axis 2: the x-y plane at padded k, float[ny][nx]. The slice and
 * free-surface outputs of data2D_XYZ_out (data_io.cpp:485-587, pack_iodata_x/y/z
 * :96-245) read these planes. */
int cgfdm_get_slice(cgfdm_solver* s, int comp, int axis, int index, float* out);
/* Overwrite one wavefield component (initial conditions; no reference counterpart). */
int cgfdm_set_field(cgfdm_solver* s, int comp, const double* in);

/* Setup products for inspection: what = 0..8 metrics, 9 J, 10 mu, 11 lambda, 12 1/rho
 * (double[NZ][NY][NX]); 20..28 rDZ_DX, 30..38 rDZ_DY (double[NY][NX]); 40 + 3*axis +
 * {0 alpha, 1 1/beta, 2 d/beta} (double[NX|NY|NZ]). (CJM.cpp:107-147, init_pml_para.cpp)
 * AWENO: what = 0..18, the SCFDM CJM planes (0..8 J*xi_x..J*zt_z, 9 J, 10..12 medium,
 * 13..18 x_xi, y_xi, z_xi, x_et, y_et, z_et), double[NZ][NY][NX]. */
int cgfdm_get_setup(cgfdm_solver* s, int what, double* out);
/* [Hmin, Hmax, Vs_min, Vs_max, Vp_min, Vp_max, rho_min, rho_max] (calc_CFL.cpp:316-340). */
int cgfdm_get_ranges(cgfdm_solver* s, double* out8);

/* Adjoint (CGFDM scheme, coef_storage 0). The reference has no adjoint; these serve the
 * waveform-gradient workflow (gradients of a misfit of the station velocities with respect to the
 * medium and the source table).
 * cgfdm_forward_record: run nsteps steps from the current state (station records, PGV and
 * displacement update as cgfdm_step does), keep a checkpoint of the state (W and PML
 * auxiliaries) every checkpoint_every steps, and return the station velocity
 * rec[ns][nsteps][3] in the wavefield precision. */
int cgfdm_forward_record(cgfdm_solver* s, int nsteps, int checkpoint_every, double* rec);
/* cgfdm_adjoint: discrete adjoint of the last cgfdm_forward_record for the adjoint source
 * src[ns][nsteps][3] = dJ/d rec. Segments between checkpoints are recomputed (the last one,
 * and each segment's RK4 stage states, are kept in device memory when they fit the budget,
 * else the stages of every step are recomputed from its start), and the transposed stages
 * run backwards. Outputs (any may be NULL): dJ/dmu, dJ/dlambda, dJ/d(1/rho) per padded cell
 * (double[NZ][NY][NX]), and dJ/d(table given to cgfdm_set_source) as double[npts][nt][6].
 * The solver is left at the end of the forward run. The checkpoints, the stage states and
 * the adjoint's work buffers stay allocated until cgfdm_destroy (reused, and grown when
 * needed, by the next cgfdm_forward_record). */
int cgfdm_adjoint(cgfdm_solver* s, const double* src, double* gmu, double* glam, double* gb, double* grate);

/* Per-kernel timing with HIP events on the solver stream. enable = 1 starts
 * accumulation (and clears it), 0 stops. out[0..5] = {interior ms, interior
 * launches, surface ms, surface launches, step ms, steps}; out may be NULL.
 * AWENO: {face-flux ms, launches, update ms, launches, step ms, steps}. */
int cgfdm_timing(cgfdm_solver* s, int enable, double* out);

/* CFL inputs without a solver: [Hmin, Hmax, Vs_min, Vs_max, Vp_min, Vp_max, rho_min,
 * rho_max] of a grid, FP64 on the GPU (calc_CFL.cpp:56-313), for choosing DT. */
int cgfdm_grid_ranges(int nx, int ny, int nz, const float* coord, const float* medium, double* out8);

/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): cgfdm_grid_ranges on a grid already in device memory, allocating
 * nothing. coord and medium are device pointers (float[3][NZ][NY][NX] each, the
 * layout of cgfdm_grid_ranges), out8 is a device pointer to 8 doubles that receives
 * [Hmin, Hmax, Vs_min, Vs_max, Vp_min, Vp_max, rho_min, rho_max]; the work is
 * enqueued on `stream` (hipStream_t as void*, NULL = default stream) and the call
 * returns without waiting. The per-block partials live in a 64 KiB device array of
 * the library that each GPU holds until the process ends and that every call on
 * that GPU shares: calls that overlap on one GPU race. Limits: nx, ny, nz in
 * 1 .. 2^30 and at most 2^40 padded cells, else CGFDM_ERR_ARG. */
int cgfdm_grid_ranges_noalloc(int nx, int ny, int nz, const float* coord, const float* medium, double* out8,
                              void* stream);

/* Name of the kernel the solver launches (no reference counterpart; for profiles
 * and reports): which = 0 the interior stage kernel (CGFDM: cg_cell, cg_zreg, ...)
 * or the face-flux kernel (AWENO), 1 the surface kernel (CGFDM) or the update
 * kernel (AWENO). NULL for a bad handle or index. */
const char* cgfdm_kernel_name(cgfdm_solver* s, int which);

/* Device bytes held by the solver. */
long long cgfdm_device_bytes(cgfdm_solver* s);

/* Last error message of this thread. */
const char* cgfdm_last_error(void);

int cgfdm_destroy(cgfdm_solver* s);

#ifdef __cplusplus
}
#endif
#endif
