/*
 * cleo_hip.h -- C ABI of libcleo_hip.so, the MI355X (gfx950) super-droplet hot path.
 *
 * Every entry point is plain C: pointers, sizes and POD structs, no C++ or torch types.
 * A context owns all device memory for one domain on one GPU and one HIP stream.
 * Calls enqueue work on that stream; calls that return data to the host synchronise it.
 * Return value of every int function: 0 on success, a negative CLEO_E_* code otherwise
 * (cleo_last_error_message gives the text).  Device-side failures that the reference
 * turns into Kokkos::abort are recorded in a sticky device error word (cleo_get_error)
 * and make the next synchronising call return CLEO_E_DEVICE.
 *
 * Units follow the reference: drop radius in units of R0 = 1e-6 m, solute mass in MASS0,
 * coordinates in COORD0 = 1000 m, gridbox volume in VOL0 = 1e9 m^3, temperature in
 * TEMP0 = 273.15 K, pressure in P0 = 1e5 Pa (libs/cleoconstants.hpp:65-104).  Times are
 * integer model steps, 1 s = 100 steps (libs/initialise/timesteps.hpp:34-59).
 *
 * Gridbox indexing is the reference's Cartesian one: gbxindex = iz + nz*(ix + nx*iy)
 * (libs/cartesiandomain/doubly_periodic_domain.hpp:41-70), out-of-domain = 0xFFFFFFFF
 * (libs/cleoconstants.hpp:115).
 */
#ifndef CLEO_HIP_H_
#define CLEO_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLEO_ABI_VERSION 1
#define CLEO_OOB_GBXINDEX 0xFFFFFFFFu

/* status codes */
enum {
  CLEO_OK = 0,
  CLEO_E_ARG = -1,     /* bad argument (null pointer, size beyond capacity, ...) */
  CLEO_E_HIP = -2,     /* HIP runtime call failed */
  CLEO_E_DEVICE = -3,  /* a kernel recorded an error in the device error word */
  CLEO_E_STATE = -4,   /* call out of order (e.g. collisions with an unconsumed/absent RNG) */
  CLEO_E_NOMEM = -5,
  CLEO_E_FILE = -6     /* a file could not be read or does not follow the expected layout */
};

/* device error codes: one per Kokkos::abort site of the reference on this path */
enum {
  CLEO_ERR_NONE = 0,
  CLEO_ERR_XI_ZERO = 1,          /* superdrop_attrs.hpp:156-160 set_xi(0) */
  CLEO_ERR_RADIUS_BELOW_DRY = 2, /* superdrop_attrs.hpp:175-181 set_radius */
  CLEO_ERR_NEG_CONDENSATE = 3,   /* superdrop_attrs.hpp:207-213 condensate_mass */
  CLEO_ERR_NEWTON_MAXITERS = 4,  /* impliciteuler.cpp:213-217 */
  CLEO_ERR_TEMP_NONPOSITIVE = 5, /* thermodynamic_equations.cpp:40-42 */
  CLEO_ERR_COAL_UNDEFINED = 6,   /* coalescence.cpp:108-112,131-143,177-187 */
  CLEO_ERR_BREAKUP_NFRAGS = 7,   /* breakup.hpp:199-255 */
  CLEO_ERR_CFL = 8,              /* cfl_criteria.hpp:52-55 */
  CLEO_ERR_BOUNDS = 9,           /* cartesian_motion.hpp:44-52 */
  CLEO_ERR_SORT_NONLOCAL = 10,   /* a key left the 27-neighbourhood: neighbour sort not applicable */
  CLEO_ERR_SUBSTEP = 11,         /* impliciteuler.hpp:336-339 delt < minsubdelt */
  CLEO_ERR_INJECTED_INDEX = 12,  /* an injected Fisher-Yates index j > iter */
  CLEO_ERR_XI_OVERFLOW = 13      /* massmoments_observer.cpp:31-35 xi >= uint64 max */
};

typedef struct cleo_ctx cleo_ctx;

/* Rectilinear Cartesian domain (libs/cartesiandomain/cartesianmaps.hpp:49-303).
   Face coordinates per axis, dimensionless, strictly increasing.
   periodic[axis]: 0 finite, 1 periodic; axis order {z, x, y} (coord3, coord1, coord2).
   DoublyPeriodicDomain is {0, 1, 1}. */
typedef struct {
  uint32_t ndims[3]; /* nz, nx, ny */
  const double* zhalf; /* nz+1 */
  const double* xhalf; /* nx+1 */
  const double* yhalf; /* ny+1 */
  int32_t periodic[3];
} cleo_domain_desc;

/* Structure-of-arrays superdroplet view (host or device pointers depending on the call).
   Field meaning is the reference's Superdrop (libs/superdrops/superdrop.hpp:36-60). */
typedef struct {
  uint32_t* sdgbxindex;
  double* coord3;
  double* coord1;
  double* coord2;
  uint64_t* xi;
  double* radius;
  double* msol;
  uint64_t* sdid;
} cleo_supers_soa;

/* Gridbox State, structure of arrays, one entry per gridbox (libs/superdrops/state.hpp:33-60).
   wvel/uvel/vvel are {lower face, upper face} pairs. */
typedef struct {
  double* volume;
  double* press;
  double* temp;
  double* qvap;
  double* qcond;
  double* wvel_lo;
  double* wvel_hi;
  double* uvel_lo;
  double* uvel_hi;
  double* vvel_lo;
  double* vvel_hi;
} cleo_gbx_state;

/* Solute properties; the reference fixes NaCl at compile time
   (libs/superdrops/superdrop_attrs.hpp:35-56, libs/cleoconstants.hpp:53-56). */
typedef struct {
  double rho_sol; /* dimensionless density (RHO_SOL / RHO0) */
  double mr_sol;  /* dimensionless molar mass (MR_SOL / MR0) */
  double ionic;   /* van't Hoff factor */
} cleo_solute;

/* Collision kernels (PairProbability, libs/superdrops/collisions/collisions.hpp:49-52) */
enum {
  CLEO_PROB_GOLOVIN = 0, /* golovinprob.cpp:32-43 */
  CLEO_PROB_LONG = 1,    /* longhydroprob.cpp:30-48 + hydrodynamicprob.hpp:50-66 */
  CLEO_PROB_CONST = 2,   /* constprob.hpp:32-50 */
  CLEO_PROB_LOWLIST_COAL = 3, /* LowListCoalProb<prob_tv>: Long (coaleff 1) x Low-List coaleff,
                                 lowlistprob.hpp:40-123 */
  CLEO_PROB_LOWLIST_BU = 4    /* LowListBuProb<prob_tv>: Long (coaleff 1) x (1 - Low-List coaleff),
                                 lowlistprob.hpp:136-153 */
};
/* Collision enactments (PairEnactX, collisions.hpp:64-67) */
enum {
  CLEO_ENACT_COAL = 0,     /* DoCoalescence, coalescence.cpp:39-196 */
  CLEO_ENACT_BREAKUP = 1,  /* DoBreakup<NFragments per nfrags_kind>, breakup.hpp:140-268 */
  CLEO_ENACT_COALBURE = 2  /* DoCoalBuRe<NFragments per nfrags_kind, CoalBuReFlag per flag_kind>,
                              coalbure.hpp:207-256 */
};
/* Fragment numbers of a breakup (NFragments, breakup_nfrags.hpp:30-91) */
enum {
  CLEO_NFRAGS_CONST = 0, /* ConstNFrags(nfrags), breakup_nfrags.hpp:38-48 */
  CLEO_NFRAGS_CKE = 1    /* CollisionKineticEnergyNFrags<nfrags_tv>, breakup_nfrags.hpp:50-91 */
};
/* Coalescence / breakup / rebound decision of DoCoalBuRe (CoalBuReFlag, coalbure_flag.hpp:34-198):
   1 coalescence, 2 breakup, else rebound (nothing changes) */
enum {
  CLEO_FLAG_CONST = 0,  /* ConstCoalBuReFlag(coalbure_coaleff), coalbure_flag.hpp:179-198 */
  CLEO_FLAG_SU = 1,     /* SUCoalBuReFlag<flag_tv> (Szakall and Urbich 2018), coalbure_flag.hpp:41-74 */
  CLEO_FLAG_TS = 2,     /* TSCoalBuReFlag<flag_tv> (Testik 2011 + Straub 2010), coalbure_flag.hpp:76-147 */
  CLEO_FLAG_STRAUB = 3  /* StraubCoalBuReFlag<flag_tv> (Straub 2010), coalbure_flag.hpp:149-177 */
};
/* Terminal velocity formulas (VelocityFormula, terminalvelocity.hpp:36-124) */
enum {
  CLEO_TV_NULL = 0,
  CLEO_TV_SIMMEL = 1,
  CLEO_TV_ROGERS_YAU = 2,
  CLEO_TV_ROGERS_GK = 3
};

typedef struct {
  uint32_t interval;    /* collstep [model steps]; DELT = step2realtime(interval) */
  int32_t prob_kind;    /* CLEO_PROB_* */
  double prob_param;    /* CONST: kernel [m^3/s]; LONG: coaleff (1.0 default); GOLOVIN: unused */
  int32_t enact_kind;   /* CLEO_ENACT_* */
  double nfrags;        /* ConstNFrags (breakup / coalbure) */
  double coalbure_coaleff; /* ConstCoalBuReFlag coaleff */
  int32_t nfrags_kind;  /* CLEO_NFRAGS_* (breakup / coalbure); 0 = ConstNFrags(nfrags) */
  int32_t flag_kind;    /* CLEO_FLAG_* (coalbure); 0 = ConstCoalBuReFlag(coalbure_coaleff) */
  int32_t prob_tv;      /* CLEO_TV_* of the Low-List probabilities (their template argument) */
  int32_t nfrags_tv;    /* CLEO_TV_* of CLEO_NFRAGS_CKE */
  int32_t flag_tv;      /* CLEO_TV_* of the SU / TS / Straub flags */
} cleo_coll_params;

typedef struct {
  uint32_t interval;     /* condstep [model steps]; delt = step2dimlesstime(interval) */
  int32_t do_alter_thermo;
  uint32_t maxniters;
  double rtol;
  double atol;
  double minsubdelt_s;   /* MINSUBTSTEP [s] */
  int32_t warm_start;    /* 0 plain guess (reference); 1 MLP guess (see cleo_mlp_*) */
  uint32_t warm_restart_iters; /* MLP path: iterations before restarting from the plain guess */
} cleo_cond_params;

typedef struct {
  uint32_t interval;     /* motionstep [model steps]; delt = step2dimlesstime(interval) */
  int32_t terminalv;     /* CLEO_TV_* */
} cleo_motion_params;

/* SDMMethods configuration with the microphysics Condensation >> Collisions plus motion
   (libs/runcleo/sdmmethods.hpp:110-333). enable_* = 0 gives the reference's Null* types.
   cleo_set_sdm(p) is cleo_set_sdm_chain with the enabled ones of {cond, coll} in that order. */
typedef struct {
  int32_t enable_cond;
  int32_t enable_colls;
  int32_t enable_motion;
  cleo_cond_params cond;
  cleo_coll_params coll;
  cleo_motion_params motion;
} cleo_sdm_params;

/* One member of a chain of microphysical processes (MicrophysicalProcess,
   libs/superdrops/microphysicalprocess.hpp:31-53): a ConstTstepMicrophysics of condensation
   (Condensation, condensation.hpp:141-150) or of collisions (CollCoal / CollBu / CoalBuRe,
   coalescence.hpp, breakup.hpp, coalbure.hpp), with its own interval in its params. */
enum {
  CLEO_PROCESS_CONDENSATION = 0,
  CLEO_PROCESS_COLLISIONS = 1
};
#define CLEO_MAX_PROCESSES 8
typedef struct {
  int32_t kind;           /* CLEO_PROCESS_* */
  cleo_cond_params cond;  /* read when kind is CLEO_PROCESS_CONDENSATION */
  cleo_coll_params coll;  /* read when kind is CLEO_PROCESS_COLLISIONS */
} cleo_process;

typedef struct {
  int32_t code;      /* CLEO_ERR_* of the first failure */
  uint32_t gbxindex; /* gridbox of the first failure */
  uint64_t drop;     /* array position of the drop involved */
} cleo_error;

/* Cumulative counters since the last cleo_reset_stats. */
typedef struct {
  uint64_t cond_drops;       /* drop-solves */
  uint64_t newton_iters;     /* calls of the Newton update (impliciteuler.cpp:240-255) */
  uint64_t substep_drops;    /* solves that took the adaptive sub-stepping branch */
  uint64_t warm_restarts;    /* MLP-guess solves that fell back to the plain guess */
  uint64_t coll_pairs;       /* pairs considered */
  uint64_t coll_events;      /* pairs with gamma != 0 */
  uint64_t nulls;            /* superdroplets set null by collisions */
  uint64_t sorts;            /* global sorts */
  uint64_t motion_drops;     /* drops moved */
} cleo_stats;

/* ---- context ---------------------------------------------------------------------- */
/* Replaces SupersInDomain + gridbox DualView construction (libs/gridboxes/supersindomain.hpp:55-62,
   libs/runcleo/creategbxs.hpp:51-149). maxnsupers = capacity of the superdroplet store. */
int cleo_create(cleo_ctx** ctx, uint64_t maxnsupers, const cleo_domain_desc* dom);
int cleo_destroy(cleo_ctx* ctx);
const char* cleo_last_error_message(const cleo_ctx* ctx);
int cleo_device_info(char* buf, size_t len);

/* ---- data movement ----------------------------------------------------------------- */
/* Host SoA in, stable sort by sdgbxindex, gridbox refs set (supersindomain.hpp:55-62). */
int cleo_upload_supers(cleo_ctx* ctx, const cleo_supers_soa* host, uint64_t ntot);
/* Same from the reference's 72-byte Superdrop layout (superdrop.hpp:36-60). */
int cleo_upload_supers_aos(cleo_ctx* ctx, const void* aos72, uint64_t ntot);
/* Same from device-resident SoA arrays (any device pointers on the context's GPU). The
   caller finishes the work that produced them (e.g. synchronises its own stream) first. */
int cleo_upload_supers_device(cleo_ctx* ctx, const cleo_supers_soa* dev, uint64_t ntot);
/* Whole store in reference order: in-domain drops sorted by gridbox, then out-of-domain. */
int cleo_download_supers(cleo_ctx* ctx, cleo_supers_soa* host, uint64_t* ntot, uint64_t* ndomain);
int cleo_download_supers_aos(cleo_ctx* ctx, void* aos72, uint64_t* ntot, uint64_t* ndomain);
/* refs: ngbx+1 entries, gridbox g holds [refs[g], refs[g+1]) (supersingbx.hpp:32-102). */
int cleo_download_refs(cleo_ctx* ctx, uint32_t* refs);
int cleo_upload_gbx_state(cleo_ctx* ctx, const cleo_gbx_state* host);
int cleo_download_gbx_state(cleo_ctx* ctx, cleo_gbx_state* host);
int cleo_set_solute(cleo_ctx* ctx, const cleo_solute* solute);

/* ---- random numbers ----------------------------------------------------------------- */
/* Counter mode: Philox4x32-10 keyed by seed, counter (item, gbxindex, event, stream).
   Replaces Kokkos::Random_XorShift64_Pool (collisions.hpp:292-302). */
int cleo_set_rng_counter(cleo_ctx* ctx, uint64_t seed, uint64_t first_event);
/* Injected mode, consumed by the next collision event: per gridbox g (in order) with n_g
   active drops, n_g-1 Fisher-Yates indices j (for iter = n_g-1 down to 1, 0 <= j <= iter),
   and 2*floor(n_g/2) doubles (phi_coll, phi_out per pair).  Gridboxes with n_g < 2 take none.
   (shuffle.hpp:80-87, collisions.hpp:152-155) */
int cleo_set_rng_injected(cleo_ctx* ctx, const uint32_t* shuffle_j, uint64_t nshuffle,
                          const double* phis, uint64_t nphis);

/* ---- hot path ----------------------------------------------------------------------- */
int cleo_set_sdm(cleo_ctx* ctx, const cleo_sdm_params* params);
/* SDMMethods with the microphysics procs[0] >> procs[1] >> ... >> procs[nprocs-1]
   (CombinedMicrophysicalProcess and operator>>, microphysicalprocess.hpp:65-134: on each
   sub-step every process that is on step runs, in chain order, on every gridbox; the next
   sub-step is the earliest next step of any process) and motion (NULL: NullMotion).  Any order
   and any number of processes up to CLEO_MAX_PROCESSES; nprocs = 0 is NullMicrophysicalProcess.
   Each collision process is its own collision event: it draws from its own counter-RNG event
   (the events advance in chain order), and an injected stream covers one event, so an injected
   chain is run process by process with cleo_collisions. */
int cleo_set_sdm_chain(cleo_ctx* ctx, const cleo_process* procs, uint32_t nprocs,
                       const cleo_motion_params* motion);
/* SortSupersBySdgbxindex + set_gridboxes_refs (sortsupers.hpp:93-226, movesupersindomain.hpp:170-175) */
int cleo_sort(cleo_ctx* ctx);
/* MoveSupersInDomain::run_step body with NullBoundaryConditions (movesupersindomain.hpp:234-301) */
int cleo_motion(cleo_ctx* ctx, const cleo_motion_params* params);
/* DoCondensation on every gridbox (condensation.hpp:141-150, condensation.cpp:44-137) */
int cleo_condensation(cleo_ctx* ctx, const cleo_cond_params* params);
/* DoCollisions on every gridbox, then the global re-sort if nulls appeared
   (collisions.hpp:263-278, sdmmethods.hpp:213-230) */
int cleo_collisions(cleo_ctx* ctx, const cleo_coll_params* params);
/* SDMMethods::run_step(t_mdl, t_mdl_next) (sdmmethods.hpp:319-332) with the cleo_set_sdm config. */
int cleo_run_step(cleo_ctx* ctx, uint32_t t_mdl, uint32_t t_mdl_next);
/* Same, captured once as a hipGraph for [t_mdl, t_mdl+span) and replayed nsteps times with
   t advancing by span each replay (RNG events advance on the device). Requires counter RNG. */
int cleo_run_steps_graph(cleo_ctx* ctx, uint32_t t_mdl, uint32_t span, uint32_t nsteps);

int cleo_synchronize(cleo_ctx* ctx);
/* HIP events on the context's stream (slots 0..15), for timing sequences of calls. */
int cleo_event_record(cleo_ctx* ctx, int32_t slot);
int cleo_event_elapsed(cleo_ctx* ctx, int32_t slot_begin, int32_t slot_end, float* ms);
/* number of in-domain superdroplets (refs[ngbx]) */
int cleo_ndomain(cleo_ctx* ctx, uint64_t* ndomain);
int cleo_get_error(cleo_ctx* ctx, cleo_error* err);
int cleo_clear_error(cleo_ctx* ctx);
int cleo_get_stats(cleo_ctx* ctx, cleo_stats* stats);
int cleo_reset_stats(cleo_ctx* ctx);
/* Record per-drop Newton iteration counts of the next condensation calls (array position order,
   0 for drops not solved). */
int cleo_record_niters(cleo_ctx* ctx, int32_t enable);
int cleo_download_niters(cleo_ctx* ctx, uint32_t* niters, uint64_t n);
/* Record the learned guesses (z = r^2 per drop, array position order) of the next warm-started
   condensation calls. */
int cleo_record_guess(cleo_ctx* ctx, int32_t enable);
int cleo_download_guess(cleo_ctx* ctx, double* z, uint64_t n);
/* The device's thermodynamic functions at n gridbox states (dimensionless, host arrays):
   saturation_pressure (thermodynamic_equations.cpp:39-52), supersaturation_ratio
   (thermodynamic_equations.hpp:79-90) and diffusion_factor (thermodynamic_equations.cpp:94-114), as
   the condensation evaluates them per gridbox. */
int cleo_thermo_eval(cleo_ctx* ctx, uint64_t n, const double* temp, const double* press, const double* qvap,
                     double* psat, double* s_ratio, double* ffactor);

/* ---- without allocation ---------------------------------------------------------------- */
/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).  The sort that
   cleo_upload_supers_device applies (stable by sdgbxindex, keys >= ngbx out of the domain after
   every gridbox; SortSupersBySdgbxindex + set_gridboxes_refs, sortsupers.hpp:93-226), from the
   device SoA `in` to the device SoA `out` (n drops each, not overlapping), with refs (device, ngbx+1
   entries) the first position of each gridbox and refs[ngbx] the in-domain count, enqueued on
   `stream` (a hipStream_t; NULL is the null stream) without synchronising.  No context, no
   hipMalloc: the scratch is a fixed device array of the library (about 17 MB per GPU, held until the
   process ends), so n <= CLEO_NOALLOC_MAX_SUPERS (CLEO_E_ARG beyond).  Calls that overlap on one GPU
   share that array and race; calls in one stream order themselves. */
#define CLEO_NOALLOC_MAX_SUPERS (1u << 20)
int cleo_sort_supers_noalloc(const cleo_supers_soa* in, cleo_supers_soa* out, uint32_t* refs, uint64_t n,
                             uint32_t ngbx, void* stream);

/* ---- CLEO binary initialisation files (host only: no context, no GPU) ---------------------- */
/* The layout of initialise/readbinary.cpp:22-99 (written by cleopy/writebinary.py): 4 uint32
   (first data byte, bytes of the global metadata string, number of variables, bytes of metadata
   per variable), the string, then per variable 3 uint32 (first byte, bytes per datapoint,
   datapoints), 2 chars (type, units) and a double (scale factor), little-endian and unpadded.

   Superdroplet initial conditions (init_all_supers_from_binary.cpp:57-100): variables 0..6 are
   sdgbxindex (uint32), xi (uint64), radius, msol, coord3, coord1, coord2 (double, dimensionless);
   a coordinate variable with no datapoints (fewer space dimensions) reads as 0.  sdIds are 0..n-1
   in file order (Superdrop::IDType::Gen::next).  With out == NULL only *n is set; otherwise the
   host arrays of `out` (except sdid, filled too) must hold `cap` >= n entries.  The drops go to
   the GPU with cleo_upload_supers. */
int cleo_read_initsupers(const char* path, cleo_supers_soa* out, uint64_t cap, uint64_t* n);
/* InitSupersFromBinary::fetch_data on one process (init_supers_from_binary.hpp:66-95,
   init_supers_from_binary.cpp:44-63): the file's drops (as cleo_read_initsupers), then
   maxnsupers - n uninitialised drops, out of the domain (CLEO_OOB_GBXINDEX) with NaN radius, solute
   mass and coordinates and xi = sdId = 0, so that the host arrays of `out` (maxnsupers entries each)
   hold exactly maxnsupers drops, as the reference's totsupers does.  *ninit = n.  CLEO_E_ARG when
   the file holds more than maxnsupers drops.  Upload the result with cleo_upload_supers(ctx, out,
   maxnsupers). */
int cleo_read_initsupers_padded(const char* path, uint64_t maxnsupers, cleo_supers_soa* out, uint64_t* ninit);
/* Gridbox boundaries (gbx_bounds_from_binary.cpp:29-58): variable 0 ndims (3 uint64: nz, nx, ny),
   1 gridbox indices (uint32), 2 six bounds per gridbox [zlow, zup, xlow, xup, ylow, yup]
   (double, dimensionless).  Writes ndims and, when the arrays are non-null, the face coordinates
   zhalf[nz+1], xhalf[nx+1], yhalf[ny+1] of the Cartesian grid that cleo_domain_desc takes
   (gridbox index iz + nz (ix + nx iy)); CLEO_E_FILE when a gridbox's bounds are not those of that
   grid. */
int cleo_read_gridfile(const char* path, uint32_t ndims[3], double* zhalf, double* xhalf, double* yhalf);

/* ---- coupling ------------------------------------------------------------------------- */
/* One coupling step of prescribed dynamics into every gridbox's State
   (FromFileComms::receive_dynamics / update_gridbox_state, coupldyn_fromfile/fromfilecomms.cpp:
   36-62; face indexing of fromfile_cartesian_dynamics.cpp:150-215).  Arrays of this step, in
   device memory for cleo_receive_dynamics: press, temp, qvap, qcond per gridbox (ngbx); the wind
   components on the faces of the C grid, wvel on z faces ((nz+1) nx ny values, index
   k + (nz+1)(i + nx j)), uvel on x faces (nz (nx+1) ny, k + nz (i + (nx+1) j)) and vvel on y
   faces (nz nx (ny+1), k + nz (i + nx j)), for gridbox k + nz (i + nx j).  A null wind array
   gives that component 0 on both faces (a model with fewer space dimensions).  Volumes stay. */
typedef struct {
  const double* press;
  const double* temp;
  const double* qvap;
  const double* qcond;
  const double* wvel_zfaces;
  const double* uvel_xfaces;
  const double* vvel_yfaces;
} cleo_dynamics_step;
int cleo_receive_dynamics(cleo_ctx* ctx, const cleo_dynamics_step* dev);
/* A thermodynamics file of CLEO's from-file dynamics (thermodynamicvar_from_binary,
   fromfile_cartesian_dynamics.cpp:22-31): the doubles of the file's first variable, all
   timesteps one after another.  With out == NULL only *n is set.  Host only. */
int cleo_read_thermofile(const char* path, double* out, uint64_t cap, uint64_t* n);

/* ---- observers ----------------------------------------------------------------------- */
/* 0th, 1st and 2nd moments of the droplet mass distribution of every gridbox
   (MassMomentsObserver: observers/massmoments_observer.cpp:23-73): mom0[g] = sum of xi,
   mom1[g] = sum of (float)(xi * m), mom2[g] = sum of (float)(xi * m * m), summed in float in the
   order of the gridbox's drops, with m the drop mass (superdrop_attrs.cpp:55-62).  rain != 0:
   raindrops only, radius >= 40 micrometres (calculate_rainmassmoments).  mom0, mom1 and mom2 are
   host arrays of ngbx entries.  A multiplicity of 2^64 - 1 records CLEO_ERR_XI_OVERFLOW. */
int cleo_mass_moments(cleo_ctx* ctx, int32_t rain, uint64_t* mom0, float* mom1, float* mom2);

/* Launch tiling, chosen by the autotuner; values 0 = default. */
typedef struct {
  uint32_t cond_waves_per_block;   /* gridboxes (waves) per workgroup, 1..4 */
  uint32_t coll_waves_per_block;
  uint32_t motion_waves_per_block;
  uint32_t cond_flat;              /* 1: one thread per drop + per-gridbox prep/sum kernels;
                                      2: the same with each wave's drops fed to its lanes from a
                                         task ring (lanes do not wait for slower drops) */
  uint32_t motion_flat;            /* 1: one thread per drop (two per thread), direction counts by
                                         atomics; 2: motion and sort fused (a counting pass, then
                                         each drop moved again straight to its sorted place);
                                      3: a counting pass, then the sort's gather moves each drop
                                         as it copies it to its sorted place (with cond_flat 1 and
                                         a condensation due, the same gather also solves it) */
  uint32_t cond_grid_waves;        /* waves per launch; each loops over gridboxes (or drops) */
  uint32_t coll_grid_waves;        /*   0: one wave per gridbox (one thread per drop) */
  uint32_t motion_grid_waves;
  uint32_t coll_flat;              /* 1: 64 gridboxes per wave, pairs flattened over the lanes */
  uint32_t cond_defer;             /* cond_flat 1: Newton iterations after which a drop is set
                                      aside and solved by a second launch (0: never) */
} cleo_tiles;
int cleo_set_tiles(cleo_ctx* ctx, const cleo_tiles* tiles);
int cleo_get_tiles(cleo_ctx* ctx, cleo_tiles* tiles);
/* Time each candidate tiling on the resident population and keep the fastest. */
int cleo_autotune(cleo_ctx* ctx, cleo_tiles* chosen);
/* Mean duration, by HIP events on the context's stream, of the main kernel of one stage over
   `reps` launches with the current tiling, on a copy of the resident drops (drops and State stay
   unchanged): stage 0 condensation (the per-drop kernel), 1 collisions, 2 motion, 3 motion + sort
   + condensation with the solves in the sort's gather (the gather kernel; the path cleo_run_step
   takes when condensation follows motion with thread-per-drop variants), 4 the mass-moment
   reduction of cleo_mass_moments (all drops, device arrays only). */
int cleo_time_kernel(cleo_ctx* ctx, int32_t stage, uint32_t reps, float* mean_ms);

/* ---- learned warm start ------------------------------------------------------------- */
/* MLP 6 -> 64 -> 64 -> 1, SiLU, BF16 weights, FP32 accumulate. Inputs per drop:
   (log r, log msol, s_ratio, T, p, delt) normalised by (mean, std) held in the weights blob;
   output log r_new.  Weight blob layout (float32): see CLEO_MLP_NPARAMS comment. */
#define CLEO_MLP_IN 6
#define CLEO_MLP_HID 64
/* W1[64][6], b1[64], W2[64][64], b2[64], W3[64], b3[1], xmean[6], xstd[6], ymean, ystd */
#define CLEO_MLP_NPARAMS (64 * 6 + 64 + 64 * 64 + 64 + 64 + 1)
#define CLEO_MLP_NBLOB (CLEO_MLP_NPARAMS + 6 + 6 + 2)
int cleo_mlp_set_weights(cleo_ctx* ctx, const float* blob);
int cleo_mlp_get_weights(cleo_ctx* ctx, float* blob);
/* Generate training samples from the resident population: (x[6], y) per drop from the plain
   solver without changing the store. Returns the number written (<= cap). */
int cleo_mlp_make_samples(cleo_ctx* ctx, const cleo_cond_params* params, float* x_dev_or_null,
                          float* y_dev_or_null, uint64_t cap, uint64_t* nwritten);
/* Adam training on device-resident samples (pointers are device pointers). */
typedef struct {
  uint32_t epochs;
  uint32_t batch;
  float lr;
  float beta1;
  float beta2;
  float eps;
  uint64_t seed;
} cleo_mlp_train_params;
int cleo_mlp_train(cleo_ctx* ctx, const float* x_dev, const float* y_dev, uint64_t nsamples,
                   const cleo_mlp_train_params* params, float* final_loss);
/* Forward pass and parameter gradient of the mean squared error on a batch (device pointers;
   grad has CLEO_MLP_NPARAMS entries, host pointer). Used to check gradients against the oracle. */
int cleo_mlp_forward(cleo_ctx* ctx, const float* x_dev, float* y_dev, uint64_t n);
int cleo_mlp_loss_grad(cleo_ctx* ctx, const float* x_dev, const float* y_dev, uint64_t n,
                       float* loss, float* grad_host);

#ifdef __cplusplus
}
#endif

#endif /* CLEO_HIP_H_ */
