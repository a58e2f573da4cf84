/*
 * pmx.h -- C ABI of the MI355X (gfx950) particle-mesh kernels behind pmwd_amd.
 *
 * Plain pointers and sizes only.  Every device pointer is a HIP device
 * allocation owned by the caller (torch in pmwd_amd); the context owns only
 * its scratch.  All work is enqueued on the context's stream; nothing blocks
 * the host except the entry points documented as returning host values.
 * Return value: 0 on success, a PMX_E* code otherwise; pmx_last_error()
 * returns the message of the last failure on the calling thread.
 *
 * Layouts (pmwd's): pmid (n,3) int16 row-major, disp/vel/acc (n,3) float32
 * row-major, meshes float32 C-order (n0,n1,n2) unless a "padded" layout is
 * named, complex meshes interleaved float32 pairs (n0,n1,n2/2+1).
 *
 * Each declaration names the reference interface it replaces; paths are in
 * eelregit/pmwd.
 */
#ifndef PMX_H
#define PMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PMX_OK = 0,
  PMX_EINVAL = 1,   /* bad argument (shape, size, null pointer)          */
  PMX_ENOMEM = 2,   /* device allocation failed                         */
  PMX_EHIP = 3,     /* HIP runtime or launch error                      */
  PMX_EUNSUP = 4    /* valid request outside what the kernels implement  */
};

typedef struct pmx_ctx pmx_ctx;

/* Library version (major*10000 + minor*100 + patch). */
int pmx_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* pmx_last_error(void);

/* ------------------------------------------------------------ context ---
 * One context per (mesh shape, cell size, device); it holds FFT twiddles,
 * k tables, three complex mesh buffers and the scatter binning scratch.
 * Replaces the static `Configuration` jit aux data (pmwd/configuration.py:21)
 * and the XLA buffers of gravity (pmwd/gravity.py:47-72).                   */
int pmx_ctx_create(int n0, int n1, int n2, double cell_size, void* stream,
                   pmx_ctx** out);
int pmx_ctx_destroy(pmx_ctx* ctx);
/* Rebinds the context to `stream`; work enqueued from now on runs on it,
 * after everything the context had queued on its previous stream (unless
 * either stream is under graph capture: then the caller orders them).    */
int pmx_ctx_set_stream(pmx_ctx* ctx, void* stream);
/* Device bytes held by the context (host value). */
long long pmx_ctx_bytes(const pmx_ctx* ctx);
/* Frees context scratch once the context's stream has finished; the next
 * call reallocates what it needs.  what = PMX_TRIM_PIPELINE: the particle
 * copy of pmx_nbody_steps (32 B per record slot); PMX_TRIM_ALL: every
 * particle-sized buffer (records, indices, value channels, that copy) and
 * extra halo channels.  No counterpart in pmwd, where XLA owns every
 * buffer: for callers that need the memory between phases (an adjoint
 * after a forward, a backward recomputing its inputs, a serving pool under
 * pressure).                                                               */
#define PMX_TRIM_PIPELINE 1
#define PMX_TRIM_ALL 2
int pmx_ctx_trim(pmx_ctx* ctx, int what);

/* Device allocator for contexts created after the call (both NULL: hipMalloc
 * and hipFree, the default).  alloc returns a block of `bytes` on `device`
 * for use on `stream` (the context's stream then), or NULL, which the
 * calling entry point reports as PMX_ENOMEM; release takes it back (bytes is
 * 0 when a context is destroyed).  The context's stream, which
 * pmx_ctx_set_stream orders after its earlier streams, is synchronised
 * before a block is taken or handed back.  No counterpart in pmwd, where
 * XLA owns every buffer: for hosts with a caching allocator of their own
 * (torch's), so that memory a context gives back serves the host's arrays
 * and the reverse without new device mappings.                            */
typedef void* (*pmx_alloc_fn)(size_t bytes, int device, void* stream, void* user);
typedef void (*pmx_free_fn)(void* ptr, size_t bytes, int device, void* stream, void* user);
int pmx_set_allocator(pmx_alloc_fn alloc, pmx_free_fn release, void* user);

/* ------------------------------------------------------- cell keys/sort ---
 * key[i] = C-order raveled index (strides of Particles.raveled_id,
 * pmwd/particles.py:156-182) of particle i's lower CIC corner, computed
 * bit-exactly as enmesh (pmwd/pm_util.py:119-137).                          */
int pmx_cell_keys(pmx_ctx* ctx, const int16_t* pmid, const float* disp,
                  long long n, uint32_t* key);
/* perm = stable ascending argsort of key (n entries, keys < 2^31).  No pmwd
 * counterpart: the permutation that orders particles by cell.             */
int pmx_sort_perm(pmx_ctx* ctx, const uint32_t* key, long long n,
                  uint32_t* perm);
/* tile[i] = raveled index of the scatter tile holding the lower corner,
 * and the tile extent used by pmx_scatter_mass (host values).              */
int pmx_tile_keys(pmx_ctx* ctx, const int16_t* pmid, const float* disp,
                  long long n, uint32_t* tile);
int pmx_tile_shape(const pmx_ctx* ctx, int tile_shape[3]);

/* ------------------------------------------------------- white noise ---
 * pmwd/modes.py:12-49 (white_noise, real=True): out[i] (n float32) standard
 * normals from Philox4x32-10 keyed by seed (two 32-bit words, low first),
 * counter = (i/4 low, high, 0, 0), and the float64 Box-Muller transform of
 * its word pairs; element i is normal i%4 of its block.  out must be 16-byte
 * aligned.  pmwd draws with jax.random (threefry), whose bit stream this does
 * not reproduce.                                                            */
int pmx_white_noise(pmx_ctx* ctx, unsigned long long seed, long long n, float* out);
/* The generator's raw words: out[4b+j] = word j of block b < nblocks (for
 * checks against published known answers).  No pmwd counterpart.          */
int pmx_random_bits(pmx_ctx* ctx, unsigned long long seed, long long nblocks, uint32_t* out);

/* --------------------------------------------------------------- LPT ---
 * pmwd/lpt.py:136-212 at lpt_order = order (0, 1 or 2) on a context of the
 * particle grid (mesh = ptcl_grid_shape, cell = ptcl_spacing).  modes is the
 * linear-modes spectrum (n0,n1,n2/2+1) complex64 of modes.py:67-117, divided
 * by cell_vol here (lpt.py:164); fac (host) = {D1, a^2 H D1', D2, a^2 H D2'}
 * in float32 (lpt.py:196-201).  disp and vel (n,3), n = n0 n1 n2 in C order
 * (Particles.gen_grid), are added to order by order: += D_m g_m and
 * a^2 H D_m' g_m with g_m = irfftn(neg_grad(k, pot_m)), pot_1 =
 * laplace(modes / cell_vol), pot_2 = laplace(rfftn(L)) and L the strain
 * source of lpt.py:40-76.  Order 2 holds one more mesh in the context.     */
int pmx_lpt(pmx_ctx* ctx, const float* modes, float cell_vol, int order,
            const float* fac, float* disp, float* vel);
/* pmwd/modes.py:67-117 (linear_modes) on an r2c spectrum (n0,n1,n2/2+1)
 * complex64 in place, on a context of the particle grid: modes *= scale
 * (float32; the 'ortho' normalisation of the white-noise transform), then
 * modes *= sqrt(V P(|k|)) in float32 with |k| from the float32 wavevectors
 * and P in float64 (boltzmann.py:399-455): P = amp (k/k_pivot)^n_s
 * (q T(k))^2 d2, T linearly interpolated in the host table (tk ascending,
 * tT; ntab entries, clamped at the ends), par (host) = {amp = 0.32 A_s
 * k_pivot, k_pivot, n_s, q = pi (c/H_0)^2 / Omega_m, V, d2 = D(a)^2 or 1}. */
int pmx_linear_modes(pmx_ctx* ctx, float* modes, float scale, const double* tk,
                     const double* tT, int ntab, const double* par);

/* ----------------------------------------------------------- scatter ---
 * Default-argument CIC mass assignment (pmwd/scatter.py:33-83 with
 * mesh=None, offset=0, cell_size=None): mesh[c] (+)= sum val*w.  Particles
 * are binned by mesh tile and accumulated in LDS; no float atomics on
 * global memory.  mesh is (n0,n1,n2) C-order; accumulate=0 overwrites it.  */
int pmx_scatter_mass(pmx_ctx* ctx, const int16_t* pmid, const float* disp,
                     long long n, float val, float* mesh, int accumulate);

/* pmx_scatter_mass with no context and no device allocation.  No
 * counterpart in pmwd: for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, hosts that forbid it).  The
 * mesh (n0,n1,n2) of cell cell_size is given directly and the work is
 * enqueued on `stream` of the current device.  The binning and halo scratch
 * live in one static 64 MiB device array per device (about 1.9 million
 * particles on a 128^3 mesh); larger requests return PMX_EINVAL with a
 * message and enqueue nothing.  Every call on a device shares that device's
 * array: calls that overlap on one device race, and the caller serialises
 * them.  Bitwise the result of pmx_scatter_mass.                          */
int pmx_scatter_mass_noalloc(const int16_t* pmid, const float* disp, long long n,
                             int n0, int n1, int n2, double cell_size, float val,
                             float* mesh, int accumulate, void* stream);

/* Generic CIC geometry for non-default arguments: enmesh's a2 branch
 * (pmwd/pm_util.py:99-118) when out_cell_size > 0, offsets b12, a target
 * shape (indices outside are dropped, pmwd/gather.py:75), channels.        */
typedef struct {
  int mesh_shape[3];      /* periodic shape s1 = conf.mesh_shape           */
  double cell_size;       /* a1 = conf.cell_size                           */
  double offset[3];       /* b12                                           */
  double out_cell_size;   /* a2, or <= 0 for None                          */
  int out_shape[3];       /* s2, target mesh shape                         */
} pmx_cic_geom;

/* pmwd/scatter.py:33-83 (any arguments): mesh (out_shape + (nchan,)) +=
 * val * w.  val is (n,nchan) or NULL for the scalar val_scalar.  Uses
 * global float atomics (not the hot path).                                  */
int pmx_scatter_cic(pmx_ctx* ctx, const pmx_cic_geom* g, const int16_t* pmid,
                    const float* disp, long long n, const float* val,
                    float val_scalar, int nchan, float* mesh);
/* pmwd/gather.py:33-77: out (n,nchan) = val + sum mesh[c]*w.  val is
 * (n,nchan) or NULL for val_scalar.                                        */
int pmx_gather_cic(pmx_ctx* ctx, const pmx_cic_geom* g, const int16_t* pmid,
                   const float* disp, long long n, const float* mesh,
                   int nchan, const float* val, float val_scalar, float* out);
/* pmwd/scatter.py:86-148 (_scatter_chunk_adj/_scatter_bwd):
 * disp_cot (n,3) = sum_j (mesh_cot[c_j].val) grad w_j / cell,
 * val_cot (n,nchan) = sum_j mesh_cot[c_j] w_j (either output may be NULL). */
int pmx_scatter_cic_adj(pmx_ctx* ctx, const pmx_cic_geom* g,
                        const int16_t* pmid, const float* disp, long long n,
                        const float* mesh_cot, const float* val,
                        float val_scalar, int nchan, float* disp_cot,
                        float* val_cot);
/* pmwd/gather.py:80-142 (_gather_chunk_adj/_gather_bwd):
 * disp_cot (n,3) = sum_j (val_cot.mesh[c_j]) grad w_j / cell and
 * mesh_cot += scatter(val_cot * w) (either output may be NULL).            */
int pmx_gather_cic_adj(pmx_ctx* ctx, const pmx_cic_geom* g,
                       const int16_t* pmid, const float* disp, long long n,
                       const float* mesh, const float* val_cot, int nchan,
                       float* disp_cot, float* mesh_cot);

/* pmwd/pm_util.py:33-156 (enmesh) as called by ptcl_enmesh
 * (pmwd/particles.py:212-254): ind (n, 2^dim, dim) int16 multilinear
 * indices, frac (n, 2^dim) fractions and, when frac_grad is non-NULL, their
 * gradients (n, 2^dim, dim).  g gives the periodic shape, cell size, offset,
 * optional output cell size and target shape; wrap = 0 drops the periodic
 * wrap (enmesh's s1 = None), and then drop = 1 replaces negative indices by
 * the target extent (pm_util.py:138-141).                                   */
int pmx_enmesh(pmx_ctx* ctx, const pmx_cic_geom* g, int wrap, int drop,
               const int16_t* pmid, const float* disp, long long n, int16_t* ind,
               float* frac, float* frac_grad);
/* The three readouts of pmwd/gravity.py:60-69 (gather.py:33-77 with default
 * arguments) in one pass: acc (n,3) from the component-major force meshes
 * mesh3 (3, n0, n1, n2) of the context's mesh, bitwise the three
 * pmx_gather_cic calls.                                                     */
int pmx_gather_cic3(pmx_ctx* ctx, const int16_t* pmid, const float* disp, long long n,
                    const float* mesh3, float* acc);

/* --------------------------------------------------------------- FFT ---
 * Hand-written Stockham FFTs.  pmx_rfftn: jnp.fft.rfftn, backward norm
 * (pmwd/pm_util.py:281); pmx_irfftn: jnp.fft.irfftn(s=mesh), scaled 1/M
 * (pmwd/pm_util.py:336).  in/out may not alias.                             */
int pmx_rfftn(pmx_ctx* ctx, const float* in, float* out_c);
int pmx_irfftn(pmx_ctx* ctx, const float* in_c, float* out);

/* ---------------------------------------------------- power spectrum ---
 * The binning of pmwd/spec_util.py:47-147 (powspec) on the context's mesh:
 * the unnormalised r2c of f (and of g for a cross spectrum), per complex
 * cell |k| of fftfreq(shape, None) in float32, P = |F|^2 or F conj(G)
 * (times sinc(k_d)^-deconv per axis when deconv > 0), the mode weight N (1
 * on the kz = 0 and even-n2 Nyquist planes, else 2) and the bin among the
 * float64 edges (host array, nedges <= 2048; right = 1: right-inclusive).
 * out (host, 4*nedges doubles): per bin the float64 sums of k N, Re(P) N,
 * Im(P) N and N; cells beyond the last edge are dropped.  The caller
 * forms the averages, cuts and units as powspec does.                      */
int pmx_powspec(pmx_ctx* ctx, const float* f, const float* g, int nedges,
                const double* edges, int right, int deconv, double* out);

/* ----------------------------------------------------------- gravity ---
 * Force meshes of pmwd/gravity.py:51-67: g_d = irfftn(neg_grad(k_d,
 * laplace(k, rfftn((dens - 1) * coef)))), coef = 1.5*Omega_m in float32.
 * grad is 3 meshes (3, n0, n1, n2) C-order.                                */
int pmx_poisson_grad(pmx_ctx* ctx, const float* dens, float coef, float* grad);
/* VJP of pmx_poisson_grad with respect to dens -- the transposes jax.vjp
 * takes through pmwd/gravity.py:51-67 (irfftn, neg_grad, laplace's own VJP
 * gravity.py:19-34, rfftn, the coef scale): dens_cot = sum_d (d grad_d /
 * d dens)^T grad_cot_d.  grad_cot is 3 meshes (3, n0, n1, n2) C-order.      */
int pmx_poisson_grad_adj(pmx_ctx* ctx, const float* grad_cot, float coef, float* dens_cot);
/* pmwd/gravity.py:47-72 with default scatter/gather arguments:
 * acc (n,3) for particles of mass val = mesh_size/ptcl_num.                */
int pmx_gravity(pmx_ctx* ctx, const int16_t* pmid, const float* disp,
                long long n, float val, float coef, float* acc);

/* ------------------------------------------------------- time stepping ---
 * pmwd/nbody.py:39-77 drift/kick in float32 with factors computed in
 * float64 by the caller and rounded to float32.                             */
int pmx_kick(pmx_ctx* ctx, long long n, float* vel, const float* acc, float k);
int pmx_drift(pmx_ctx* ctx, long long n, float* disp, const float* vel, float d);
/* One KDK step of pmwd/nbody.py:121-140 (default symp_splits): vel += acc*k1;
 * disp += vel*d; acc = gravity(disp); vel += acc*k2.                       */
int pmx_nbody_step(pmx_ctx* ctx, const int16_t* pmid, float* disp, float* vel,
                   float* acc, long long n, float val, float coef, float k1,
                   float d, float k2);
/* nsteps KDK steps of pmwd/nbody.py:121-140 (the loop of nbody, :215-223)
 * with factors k1[s], d[s], k2[s] (host arrays): bitwise the result of
 * nsteps pmx_nbody_step calls.  Between steps the particles stay on the
 * device in tile order (two extra copies of about 32 bytes per particle);
 * disp, vel and acc are written once, at the end.  Meshes with an axis
 * longer than 1024 cells run as nsteps pmx_nbody_step calls.               */
int pmx_nbody_steps(pmx_ctx* ctx, const int16_t* pmid, float* disp, float* vel,
                    float* acc, long long n, float val, float coef, int nsteps,
                    const float* k1, const float* d, const float* k2);
/* pmx_nbody_steps with every factor read by the kernels from the device
 * table fac (float32, 1 + 3*nsteps entries): {1.5*Omega_m, k1[0], d[0],
 * k2[0], k1[1], d[1], k2[1], ...}; bitwise the same result.  Same loop
 * (pmwd/nbody.py:215-223); no host copy of the factors is taken, so a
 * graph that captures the call replays for whatever the table holds.  Needs
 * n > 0 and mesh axes of at most 1024 cells.                               */
int pmx_nbody_steps_dev(pmx_ctx* ctx, const int16_t* pmid, float* disp, float* vel,
                        float* acc, long long n, float val, int nsteps, const float* fac);

/* -------------------------------------------------------- graphs, packs ---
 * The N-body part of pmwd's forward model (nbody_init's force,
 * pmwd/nbody.py:193-201, the KDK loop :215-223, then scatter.py:8-83 of the
 * final state into dens when dens is not NULL) captured into a HIP graph
 * that the context keeps, factors read from the device table fac as in
 * pmx_nbody_steps_dev.  Replays (pmx_graph_launch) run on the context's
 * stream with the buffers given here and whatever fac then holds.  Run the
 * same call uncaptured once before (pmx_nbody_steps_dev and pmx_scatter_mass
 * on the same n) so that the context holds all its scratch: nothing may be
 * allocated under capture.  Timing must be off.  The graph addresses the
 * context's scratch, so freeing any of it (pmx_ctx_trim, or a later call
 * that regrows it) drops the graph: pmx_graph_launch then fails until the
 * next capture.  Packs hold their own copy and must not outlive the
 * scratch of their contexts either.  A pack launches the graphs
 * of k contexts (distinct buffers) as independent branches of one graph:
 * one launch runs k simulations side by side.  pmx_pack_create copies the
 * graphs; capturing a context again leaves existing packs unchanged.  No
 * pmwd counterpart: jit dispatch there.                                    */
/* A non-blocking stream at the default priority on the given device, for
 * a caller that captures: streams from a host's pool can be shared with
 * other users, and work they queue during a capture invalidates it.
 * pmx_stream_destroy waits for the stream, then destroys it.  No pmwd
 * counterpart.                                                              */
int pmx_stream_create(int device, void** out);
int pmx_stream_destroy(void* stream);
typedef struct pmx_pack pmx_pack;
int pmx_graph_capture(pmx_ctx* ctx, const int16_t* pmid, float* disp, float* vel,
                      float* acc, long long n, float val, int nsteps, const float* fac,
                      float* dens);
int pmx_graph_launch(pmx_ctx* ctx);
int pmx_pack_create(pmx_ctx* const* ctxs, int k, pmx_pack** out);
int pmx_pack_launch(pmx_pack* pack, void* stream);
int pmx_pack_destroy(pmx_pack* pack);

/* ------------------------------------------------------------ adjoint ---
 * force_adj of pmwd/nbody.py:108-118: acc = gravity(disp) and
 * acc_cot = VJP_disp(gravity)[vel_cot]; *dot (host) = sum(vel_cot*acc) in
 * float64, from which the Omega_m cotangent sum(vel_cot*acc)/Omega_m
 * follows (acc is proportional to Omega_m).  dot may be NULL.              */
int pmx_gravity_vjp(pmx_ctx* ctx, const int16_t* pmid, const float* disp,
                    long long n, float val, float coef, const float* vel_cot,
                    float* acc, float* acc_cot, double* dot);
/* kick_adj of pmwd/nbody.py:80-99: vel += acc*k; disp_cot -= acc_cot*k;
 * *dot (host) = sum(vel_cot*acc) in float64 (may be NULL).                 */
int pmx_kick_adj(pmx_ctx* ctx, long long n, float* vel, const float* acc,
                 float* disp_cot, const float* acc_cot, const float* vel_cot,
                 float k, double* dot);
/* drift_adj of pmwd/nbody.py:49-67: disp += vel*d; vel_cot -= disp_cot*d;
 * *dot (host) = sum(disp_cot*vel) in float64 (may be NULL).                */
int pmx_drift_adj(pmx_ctx* ctx, long long n, float* disp, const float* vel,
                  const float* disp_cot, float* vel_cot, float d, double* dot);

/* One reverse-time KDK step of pmwd/nbody.py:143-162 (default symp_splits:
 * kick_adj ka, drift_adj d, force_adj, kick_adj kb) with its closing kick
 * left pending, in place:
 *   [vel += acc*kp; disp_cot -= acc_cot*kp]  the previous step's closing
 *                                            kick_adj, skipped when kp == 0;
 *   vel += acc*ka; disp_cot -= acc_cot*ka; disp += vel*d; vel_cot -= disp_cot*d;
 *   [disp = disp_snap when non-NULL: a stored forward state; disp_snap may
 *    be disp itself, holding that state on entry, so that no second array
 *    is needed];
 *   acc = gravity(disp); acc_cot = VJP_disp(gravity)[vel_cot].
 * The caller passes this step's kb as the next call's kp, and applies the
 * last one with pmx_kick_adj.  Each kick is rounded separately, so results
 * equal the four separate calls bit for bit.  On entry acc/acc_cot hold the
 * previous force and its VJP (pmx_gravity_vjp or the previous step).
 * dots (host, may be NULL): dots[0] = sum(disp_cot*vel) of the drift
 * adjoint, dots[1] = sum(vel_cot*acc) of the new force, which is also the
 * dot of every kick adjoint that uses that force.                          */
int pmx_nbody_step_adj(pmx_ctx* ctx, const int16_t* pmid, float* disp,
                       float* vel, float* acc, float* disp_cot, float* vel_cot,
                       float* acc_cot, long long n, float val, float coef,
                       float kp, float ka, float d, const float* disp_snap,
                       double* dots);
/* nsteps reverse KDK steps of integrate_adj (pmwd/nbody.py:143-162, the
 * loop of nbody_adj :226-260): bitwise the arrays of the chained calls
 * pmx_nbody_step_adj(kp, ka[0], d[0]), pmx_nbody_step_adj(kb[0], ka[1],
 * d[1]), ..., with disp_snap (may be NULL; disp itself only when nsteps
 * is 1) given to the last; the last
 * closing kick kb[nsteps-1] stays pending for the caller, as there.
 * dots[2s], dots[2s+1] (host, 2*nsteps doubles) are step s's two dots,
 * the same sums in another grouping.  Between steps the particles and
 * their cotangents stay on the device in tile order (about 100 extra
 * bytes per particle).  Meshes with an axis longer than 1024 cells run as
 * the chained calls.                                                       */
int pmx_nbody_steps_adj(pmx_ctx* ctx, const int16_t* pmid, float* disp,
                        float* vel, float* acc, float* disp_cot, float* vel_cot,
                        float* acc_cot, long long n, float val, float coef,
                        float kp, int nsteps, const float* ka, const float* d,
                        const float* kb, const float* disp_snap, double* dots);

/* float64 dot product of two float32 arrays of length n (host value).      */
int pmx_dot(pmx_ctx* ctx, const float* a, const float* b, long long n, double* out);

/* ------------------------------------------------- host cosmology tables ---
 * Host-only float64 functions (no context, no device).
 *
 * pmx_growth_ode replaces the growth integration of pmwd/boltzmann.py:163-229
 * (odeint of pmwd/ode_util.py:60-186 on the 2LPT growth ODE): the state
 * (G1, G1', G2, G2', S1, S1', S2, S2') in ln a from y0 at lna[0], S the
 * Omega_m forward sensitivities, reported at each of the na non-decreasing
 * lna; adaptive Dormand-Prince 5(4) with the error norm over the first nerr
 * components, first step dt0 (<= 0: the Hairer-Norsett-Wanner estimate).
 * Y and Yd (na x 8, row-major) receive the state and its ln a derivative.
 *
 * pmx_tophat_var replaces the variance integral of pmwd/boltzmann.py:272-291
 * (mcfit.TophatVar there): sigma^2(R) = int k^3 P W(kR)^2 / (2 pi^2) dln k
 * over nk log-spaced k with P(k) given, Simpson weights (trapezoid for an
 * even nk), for each of nR radii.                                           */
int pmx_growth_ode(double Omega_m, double Omega_k, double w_0, double w_a,
                   const double* lna, int na, const double* y0, double rtol,
                   double atol, double dt0, int nerr, double* Y, double* Yd);
int pmx_tophat_var(const double* k, const double* P, int nk, const double* R,
                   int nR, double* out);

/* ------------------------------------------------------------- timing ---
 * Kernel timing with HIP events recorded on the context stream around each
 * launch of the hot path, grouped in PMX_NTIMER_CLASSES classes (bin, scan,
 * fill, deposit, fold, z-r2c, y-fwd, poisson-x, y-inv, z-c2r, gather,
 * vjp-gather, grad-gather, kick/drift-adjoint, other).  The fold class is
 * empty where the Poisson solve's row pass adds the halo shells itself.
 * pmx_timing_read synchronises the stream and returns, per class, the summed
 * milliseconds and the number of launches since the last read.  No pmwd
 * counterpart: measurement support.                                         */
#define PMX_NTIMER_CLASSES 15
int pmx_timing_enable(pmx_ctx* ctx, int on);
int pmx_timing_read(pmx_ctx* ctx, double* ms, long long* count);

#ifdef __cplusplus
}
#endif
#endif /* PMX_H */
