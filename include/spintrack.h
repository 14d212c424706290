/*
 * spintrack.h -- C ABI of the MI355X (gfx950) spin-tracking hot path.
 *
 * This is the drop-in boundary for the reference's simulation loop.  The
 * reference has no plugin API; its loop is the `particle` class
 * (include/particle.h:87-314) driven by mainAnalysis (src/simulation.cpp:8-97):
 *
 *   particle p(opt);            -> st_create(&opt, NULL, &h)
 *   p.initParticles();          -> st_init(h)
 *   p.runSimulation(nextTOut);  -> st_advance_to(h, nextTOut)
 *   p.outputData(log);          -> st_snapshot(h, t, pos, v, S, failure,
 *                                              n_coll, n_bounce, n_steps)
 *   ~particle()                 -> st_destroy(h)
 *
 * Conventions: every entry point returns 0 on success and a negative ST_E*
 * code on failure (never exit()s, never throws across the ABI); the message
 * of the last failure is available from st_last_error().  Per-particle
 * failures keep the reference's convention (failureState 1-4 from the
 * sanity check, negative integrator codes, NaN-poisoned t/pos/v/S;
 * src/particle.cpp:412-444, 596-612), plus one code of this library's own:
 * ST_FAIL_WATCHDOG (-9), a particle stopped after st_ext.max_events events
 * in one output interval (the reference's loop has no such bound; a lane
 * whose event loop does not end would otherwise hold the GPU).  Integrator
 * codes: DOP853 -1 (nmax), RK45 / RKF45 / quaternion / Magnus -1 (h < hmin)
 * and -2 (nmax) (src/integrator.cpp:129-131, 280-286).  All pointers passed to the st_*
 * calls are host pointers unless the name ends in _device.  One handle must
 * not be used from two host threads at once (the reference's particle
 * object is not re-entrant either).
 */
#ifndef SPINTRACK_H
#define SPINTRACK_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ST_API __attribute__((visibility("default")))

/* error codes */
#define ST_OK            0
#define ST_E_ARG        -1   /* invalid argument (null pointer, bad size, ...) */
#define ST_E_HIP        -2   /* a HIP runtime call failed */
#define ST_E_UNSUPPORTED -3  /* an option this build does not implement */
#define ST_E_STATE      -4   /* call out of order (e.g. advance before init) */

/* per-particle failureState of a particle stopped by the event watchdog */
#define ST_FAIL_WATCHDOG -9
/* default per-interval event bound (st_ext.max_events == 0) */
#define ST_DEFAULT_MAX_EVENTS 100000000u
/* largest ensemble (plain or batched) one handle holds: lanes address their
 * particle's state with 32-bit byte offsets (8 bytes per particle) */
#define ST_MAX_PARTICLES (1u << 29)

/* include/double3.h:17-21 */
typedef struct st_coords { double x, y, z; } st_coords;

/* include/options.h:10-61.  Byte-identical layout (440 bytes on x86-64 and
 * on gfx950); it is also the header of the reference binary output format
 * (src/logger.cpp:27-29). */
typedef struct st_options {
    st_coords B0, E, L, yi, Gx, Gy, Gz;
    double m, t0, tf, rtol, atol, beta, uround, safe, fac1, fac2, hmax, hmin,
           h, T, sqrtKT_m, tc, gamma, V, a, w, swapStepSize, maxPosStep,
           ioutInt, diffuse;
    uint32_t nmax, seed;
    int32_t integratorType, numParticles, numPerGPUBlock, iout;
    char dist;
    uint8_t gas_coll, gravity, fixedStepSize, keepStepSize;
    st_coords noiseAmplitudes, noiseFrequencies;
} st_options;

/* Knobs the reference hard-codes in its user-editable hooks, plus the
 * choice of arithmetic variant.  Pass NULL for the reference behaviour. */
#define ST_VARIANT_EXACT 0  /* the reference's operation order, no FMA contraction */
#define ST_VARIANT_FAST  1  /* FMA-contracted, per-segment field polynomial; agrees
                               with EXACT to the integrator tolerance */
#define ST_ENGINE_PIPELINE 0  /* kinematics kernel + spin kernel over staged segment records */
#define ST_ENGINE_FUSED    1  /* one kernel interleaving events and integrator trials */
typedef struct st_ext {
    int32_t pulse_axis;     /* axis of pulse(t) = a*cos(w*t): 0,1,2; reference = 2 (z),
                               src/integrator.cpp:40-42 */
    int32_t variant;        /* ST_VARIANT_* */
    int32_t block_size;     /* threads per workgroup of the kinematics kernel (pipeline) or of the
                               fused per-interval kernel; multiples of 64 up to 256; 0 = 256 */
    int32_t engine;         /* ST_ENGINE_* */
    int32_t seg_buffer_mb;  /* HBM budget of the pipeline's segment-record buffer; 0 = 160 GiB, at
                               most half the device memory free when the ensemble is seeded.
                               With chunk_records 0 the depth (at most 4096 records) comes from
                               it, plus overflow pages of up to a quarter as many records again
                               (at most a quarter of the budget) for particles that reach the
                               depth near the end of the run.  Trajectory-only handles
                               (integratorType -1) keep no records and ignore it */
    int32_t chunk_records;  /* trajectory segments per particle per pipeline chunk (and between
                               adjoint checkpoints); 0 = as many as seg_buffer_mb holds; an
                               explicit depth gets no overflow pages; ignored by
                               trajectory-only handles, which run each call in one chunk */
    int32_t graph_chunks;   /* pipeline engine: capture this many chunks (rounded up to even, at
                               most 64) in one HIP graph and let st_advance_to / st_run replay it
                               until every particle is done -- one host round trip per replay
                               instead of per chunk; 0 = host-driven chunk loop.  Runs with
                               checkpointing or per-kernel timing on, and integratorType 6, use the
                               host-driven loop. */
    int32_t spectrum_sum;   /* integratorType 6: which particles' covariance spectra each interval adds
                               to the run's spectrum.  0 = the reference CPU path's: its in-place tree
                               (src/particle.cpp:826-834) leaves particles 0 and 2^k < numParticles
                               in slot 0, added in that order; 1 = every particle (aggregateSpectrum,
                               src/particle.cpp:893-897), summed in a fixed order on the device */
    uint32_t max_events;    /* events one particle may take in one output interval before it is
                               stopped with failureState ST_FAIL_WATCHDOG; 0 = ST_DEFAULT_MAX_EVENTS,
                               0xffffffff = no bound (the reference's loop) */
    uint32_t adjoint_window; /* st_adjoint: checkpoint chunks replayed and walked back per adjoint
                                launch (the same gradient for any value); 0 = as many as a record
                                window of at most 64 GiB and a quarter of the free memory holds */
} st_ext;

/* include/particle.h:52-59 */
typedef struct st_rng {
    uint64_t x, y, z, w;
    double spare;
    uint8_t hasSpare;
    uint8_t pad_[7];
} st_rng;

/* Complete per-particle state (everything the reference keeps across
 * runSimulation calls, include/particle.h:147-170), for checkpoints and tests. */
typedef struct st_particle {
    st_coords S, v, v_old, pos, pos_old;
    double t, t_old, tf, next_gas, h;
    st_rng rng;
    uint64_t n_bounce, n_coll, n_steps;
    int32_t failure;
    uint32_t part_id;
    uint8_t stop, coll_type, wall_hit, pad_[5];
} st_particle;

typedef struct st_handle st_handle;

/* == particle::particle(opt) (include/particle.h:90-166): validates the
 * options and allocates the device state for opt->numParticles particles in
 * HBM.  opt->tc and opt->sqrtKT_m are used as given (the reference's
 * mainAnalysis derives them, src/simulation.cpp:9-11).  ext may be NULL. */
ST_API int st_create(const st_options* opt, const st_ext* ext, st_handle** out);

/* Batched form (no counterpart in the reference, whose particle class holds
 * one options block, include/particle.h:90-166): n_cfg configurations in one
 * handle, particles of configuration c at indices offsets[c] ..
 * offsets[c+1]-1 (cfgs[c].numParticles each, back to back).  Per
 * configuration: B0, E, Gx/Gy/Gz, gamma, pulse a and w, tc, sqrtKT_m,
 * diffuse, V, yi, dist and seed; every other option (geometry, time axis,
 * integrator control, gravity, gas_coll, step-size switches) must equal
 * configuration 0's.  Particle j of configuration c is seeded as particle j
 * of a plain run of cfgs[c], so each configuration's particles reproduce that
 * plain run bit for bit.  All calls that take a handle work on a batch; the
 * whole batch advances together and snapshot blocks cover all particles in
 * index order.  Pipeline engine only. */
ST_API int st_create_batch(const st_options* cfgs, int32_t n_cfg, const st_ext* ext, st_handle** out);

/* Configuration layout: n_cfg and the n_cfg+1 particle offsets (a plain
 * handle reports one configuration).  Either pointer may be NULL. */
ST_API int st_batch_layout(const st_handle* h, int32_t* n_cfg, uint32_t* offsets);

/* == particle::~particle (include/particle.h:168-240) */
ST_API void st_destroy(st_handle* h);

/* Message of the most recent failure on this handle (or of the last failed
 * st_create when h is NULL). */
ST_API const char* st_last_error(const st_handle* h);

/* Run all later work of this handle on the given hipStream_t (NULL = the
 * handle's own stream).  The stream must have the default priority. */
ST_API int st_set_stream(st_handle* h, void* hip_stream);

/* == particle::initParticles (include/particle.h:241-252, initParticlesCPU
 * src/particle.cpp:638-710): seeds particle i from splitmix64(seed + i). */
ST_API int st_init(st_handle* h);

/* Seeds local slot j as global particle first + j*stride, so a deterministic
 * subset of a larger ensemble can be replayed.  st_init == (0, 1). */
ST_API int st_init_subset(st_handle* h, uint32_t first, uint32_t stride);

/* == particle::runSimulation(nextTOut) (src/particle.cpp:865-876 ->
 * runSimulationGPU :525-635): advances every particle to next_t.
 * Asynchronous on the handle's stream; st_synchronize() or any copy-out
 * call waits for it. */
ST_API int st_advance_to(st_handle* h, double next_t);

/* == particle::outputData -> BinaryLogger::writeSnapshot
 * (src/particle.cpp:852-862, src/logger.cpp:14-25): copies the eight
 * snapshot arrays to host memory in the reference's layout.  Any pointer may
 * be NULL to skip that array. */
ST_API int st_snapshot(st_handle* h, double* t, st_coords* pos, st_coords* v,
                       st_coords* S, int32_t* failure, uint64_t* n_coll,
                       uint64_t* n_bounce, uint64_t* n_steps);

/* Bytes of one snapshot block in the reference file layout: numParticles *
 * (8 + 24 + 24 + 24 + 4 + 8 + 8 + 8) = 108 * numParticles. */
ST_API size_t st_snapshot_block_bytes(const st_handle* h);

/* Writes the current state as one snapshot block (the byte image
 * BinaryLogger::writeSnapshot writes) into device memory at dst. */
ST_API int st_snapshot_block_device(st_handle* h, void* dst);

/* Persistent multi-interval run: advances every particle through the
 * output times targets[0..n_targets-1] (each > the previous) in one
 * launch, with a work queue that refills lanes as particles finish.  When
 * blocks_dev is not NULL, snapshot block k (the state at targets[k]) is
 * written to blocks_dev + k*st_snapshot_block_bytes(h).  The per-particle
 * results are identical to calling st_advance_to(targets[k]) in order. */
ST_API int st_run(st_handle* h, int32_t n_targets, const double* targets,
                  void* blocks_dev);

/* No counterpart in the reference: the form of st_run for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and per-particle results
 * as st_run, but it calls no hipMalloc/hipFree: the output times are staged
 * and the kernels' work-queue and list counters kept in one static device
 * array of the library (ST_NOALLOC_MAX_TARGETS * 8 + 256 bytes per GPU, held
 * until the process ends), so n_targets is at most ST_NOALLOC_MAX_TARGETS.
 * The handle must be seeded (st_init allocates the ensemble and the
 * pipeline's record buffer) and must not have checkpointing on; a handle
 * with graph_chunks runs the host-driven chunk loop here.  Every call on one
 * GPU shares the array and the call waits for its work before returning:
 * calls on one GPU from several host threads at once race. */
#define ST_NOALLOC_MAX_TARGETS 4096
ST_API int st_run_noalloc(st_handle* h, int32_t n_targets, const double* targets, void* blocks_dev);

/* Per-particle end times (batched extension; the reference overwrites its
 * per-particle tf array, src/particle.cpp:748-749).  Particle i takes part
 * in an interval (t_prev, t_next] only while t_prev < stop_at[i], and then
 * advances to min(t_next, stop_at[i]).  NULL clears it. */
ST_API int st_set_stop_times(st_handle* h, const double* stop_at);

/* Full state copy-out / copy-in (n = numParticles records). */
ST_API int st_get_state(st_handle* h, st_particle* out);
ST_API int st_set_state(st_handle* h, const st_particle* in);

/* Work counters summed since st_init (6 entries): accepted integrator trial
 * steps, rejected trial steps, integrated segments, spin-loop trips, events
 * (collision/move/sanity passes) and kinematic records written by the
 * pipeline.  (The reference keeps no ODE step counter, src/integrator.cpp:101.) */
ST_API int st_work_counters(st_handle* h, uint64_t out[6]);

/* Per-particle ODE work since st_init (n = numParticles entries each; any
 * pointer may be NULL): accepted and rejected trial steps, and integrated
 * segments. */
ST_API int st_particle_counters(st_handle* h, uint64_t* accepted, uint64_t* rejected,
                                uint64_t* segments);

/* Segment trace of one particle (tracing/debug facility; the reference has
 * none).  particle < 0 disables.  Later advance/run calls append two
 * records per integrated segment of that particle, ST_TRACE_DOUBLES doubles
 * each: tag (1 = segment start, 2 = end), t_old, t, pos_old[3], v_old[3],
 * S[3], step size, accepted, rejected, coll_type, wall_hit, integrator
 * status, n_bounce, n_coll.  st_get_trace copies up to capacity records and
 * reports how many were recorded (at most the capacity given to
 * st_set_trace). */
#define ST_TRACE_DOUBLES 20
ST_API int st_set_trace(st_handle* h, int64_t particle, uint32_t capacity);
ST_API int st_get_trace(st_handle* h, double* out, uint32_t capacity, uint32_t* n_records);

/* Device memory for snapshot blocks and other buffers a caller hands to
 * st_run / st_snapshot_block_device, without a second HIP runtime in the
 * process; st_memcpy_d2h copies on the handle's stream and waits. */
ST_API int st_device_alloc(size_t bytes, void** ptr);
ST_API int st_device_free(void* ptr);
/* Page-locked host memory for staging snapshot blocks: st_memcpy_d2h into it
 * runs at the link's full rate, and a host thread can write one block to the
 * output file while the device computes the next interval.  (No counterpart
 * in the reference, whose logger reads the particle arrays directly,
 * src/particle.cpp:852-862.) */
ST_API int st_host_alloc(size_t bytes, void** ptr);
ST_API int st_host_free(void* ptr);
ST_API int st_memcpy_d2h(st_handle* h, void* dst, const void* src, size_t bytes);
/* Device-to-device copy queued on the handle's stream (asynchronous). */
ST_API int st_memcpy_d2d(st_handle* h, void* dst, const void* src, size_t bytes);

/* Event timing on the handle's stream: st_timer_stop records the closing
 * event, waits for it and returns the elapsed milliseconds. */
ST_API int st_timer_start(st_handle* h);
ST_API int st_timer_stop(st_handle* h, float* ms);

/* Per-kernel timing of the pipeline engine: when enabled, every later
 * advance/run records HIP events around each kinematics and spin kernel
 * launch; st_kernel_times returns the summed milliseconds and the launch
 * counts ([0] kinematics, [1] spin).  Enabling resets the sums. */
ST_API int st_kernel_timing(st_handle* h, int enable);
ST_API int st_kernel_times(st_handle* h, double ms[2], uint64_t launches[2]);

/* Differentiation (no counterpart in the reference).  With checkpointing on,
 * each later st_advance_to / st_run keeps the kinematic state of every
 * particle at each pipeline chunk start, i.e. every st_ext.chunk_records
 * trajectory segments; a new run replaces the previous run's checkpoints. */
ST_API int st_set_checkpointing(st_handle* h, int enable);

/* Adjoint pass over the last checkpointed st_run(n_targets, targets,
 * blocks_dev), which must have run without particle failures.  For a loss
 * L that depends on the spins only through the ensemble means <S>(t_k) at
 * the output times, seeds[(c*n_targets + k)*3 + i] = dL/d<S_i>(t_k) of
 * configuration c (one configuration for a plain handle).  The kernel
 * integrates u = lambda x S backwards along the replayed trajectories
 * (DOP853, the run's rtol/atol) and returns, per configuration,
 * grad[c*15 + f] = dL/dtheta_f for theta = (Gx.x, Gx.y, Gx.z, Gy.x, Gy.y,
 * Gy.z, Gz.x, Gz.y, Gz.z, E.x, E.y, E.z, B0.x, B0.y, B0.z), i.e. the
 * gradient coefficients dB_i/dx_j (row i), E and B0.  blocks_dev holds the
 * run's snapshot blocks (their S arrays give S(t_k)). */
ST_API int st_adjoint(st_handle* h, const void* blocks_dev, const double* seeds, double* grad);

/* Ensemble sums over n_blocks consecutive snapshot blocks in device memory
 * (on-device form of the reference's analysis selection,
 * benchmarks/CheckingFrequencyShiftFromEField.ipynb:29-35): for block k and
 * configuration c, out[(k*n_cfg + c)*4 + {0,1,2,3}] = number of particles
 * with failure == 0 and finite S, and the sums of Sx, Sy, Sz over them.
 * Deterministic (fixed-order partial sums).  Waits for the result. */
ST_API int st_block_sums(st_handle* h, const void* blocks_dev, int32_t n_blocks, double* out);

/* Per-particle precession phase statistics over n_blocks consecutive
 * snapshot blocks in device memory: the reference's analysis
 * (benchmarks/CheckingFrequencyShiftFromEField.ipynb:5-27,
 * scripts/nedmSim.py:151-178) as device reductions.  Particles with
 * failure == 0 and a finite phase in the last block are selected; per block
 * and configuration phi = atan2(Sz, Sy) is aligned to the left edge of the
 * first fullest bin of a 1000-bin histogram (phases <= ref - pi gain 2 pi,
 * >= ref + pi lose 2 pi), phases outside mean -+ 5 std are dropped, and
 * out[(k*n_cfg + c)*6 + {0..5}] = mean (wrapped into [-pi, pi]), std,
 * phases kept, phases selected, the reference angle, and the mean before
 * wrapping.  phi_dev (optional, n_blocks*numParticles doubles of device
 * memory) receives the selected phases (NaN where not selected).  Waits for
 * the result. */
#define ST_PHASE_STATS 6
ST_API int st_phase_stats(st_handle* h, const void* blocks_dev, int32_t n_blocks, double* out, void* phi_dev);

/* ---- Spectral path: integratorType 6 ------------------------------------
 * The reference samples the noise field gamma*(v x E/c^2 + G.r + a sin(w t))/2
 * along each trajectory segment at the multiples of the particle's h into
 * ST_NW Goertzel filters (integrateSpectrum, src/integrator.cpp:730-743;
 * SpectrumAggregator, src/floquet.cpp:111-147), sums the per-particle
 * covariance spectra every interval (src/particle.cpp:824-834) and turns the
 * total into a final Bloch vector with a Floquet-Markov master equation
 * (floquet_integrate, src/integrator.cpp:869-889).  st_create accepts
 * integratorType 6 on plain pipeline handles with w > 0; st_init computes
 * the Floquet decomposition (floquet_diagonalize, :824-867) as the
 * reference's initParticles does; every st_advance_to / st_run interval
 * clears the aggregators, samples on the GPU and adds the interval's
 * spectrum on the GPU.  The spin S is not integrated (it only turns NaN when
 * the particle stops).  Work counter 0 and the per-particle "accepted"
 * counter count samples. */
#define ST_NW 17              /* include/floquet.h:34 */
#define ST_NK 11              /* include/floquet.h:33 */
#define ST_FLOQUET_NPROP 100  /* src/integrator.cpp:827 */

/* floquetDiagonalization (include/floquet.h:99-106); quaternions (w, x, y, z) */
typedef struct st_floquet {
    double f_modes_0[4];
    double f_energies[4];
    double frequencies[ST_NW];
    double dt;                 /* period / n_prop */
    int32_t n_prop;
    int32_t pad_;
    double propagators[ST_FLOQUET_NPROP][4];
} st_floquet;

/* CovarianceSpectrum (include/floquet.h:45-62): the summed covariance
 * matrices, variance[q][r][s] = (re, im), before normalisation. */
typedef struct st_spectrum {
    double frequencies[ST_NW];
    double dt;
    uint64_t n_samples;
    double variance[ST_NW][3][3][2];
} st_spectrum;

/* == particle::cspec after the last interval (the CPU path's accumulated
 * spectrum; the reference's GPU build never fills it). */
ST_API int st_spectrum_get(st_handle* h, st_spectrum* out);
/* == particle::getSpectrumAggregators (src/particle.cpp:899-901): the
 * per-particle filter state of the last interval, 9*ST_NW doubles per
 * particle (s1[ST_NW][3], s2[ST_NW][3], imag_diag[ST_NW][3]), and the
 * per-particle sample counts.  Either pointer may be NULL. */
ST_API int st_spectrum_aggregators(st_handle* h, double* agg, uint32_t* n_samples);
/* The handle's Floquet decomposition (particle::fd). */
ST_API int st_floquet_info(st_handle* h, st_floquet* out);
/* == particle::floquetResults (src/particle.cpp:841-843; written by
 * postProcess as "b_end", :845-850): the final Bloch vector. */
ST_API int st_floquet_result(st_handle* h, double b_end[3]);

/* Host-side Floquet routines (no GPU work), for callers that hold their own
 * spectra.  == floquet_diagonalize(opt) (src/integrator.cpp:824-867). */
ST_API int st_floquet_diagonalize(const st_options* opt, const st_ext* ext, st_floquet* out);
/* == floquet_master_equation_rates (src/floquet.cpp:192-279).  power holds
 * ST_NW (re, im) spectrum values at frequencies; Delta[2][2][ST_NK],
 * X[2][2][ST_NK][2] and Gamma[2][2][ST_NK][2] are overwritten (any may be
 * NULL); Zeta[2][2][2] and Omicron[2][2][2] are added to. */
ST_API int st_floquet_rates(const double f_modes_0[4], const double f_energies[4], const double c_op[4],
                            const double* propagators, int32_t n_prop, double period, const double* frequencies,
                            const double* power, double* Delta, double* X, double* Gamma, double* Zeta,
                            double* Omicron);
/* == integrateFloquetMarkov (src/integrator.cpp:791-822), rho[2][2][2] in place. */
ST_API int st_floquet_markov(double t0, double tf, double* rho, const double* Zeta, const double* Omicron);
/* == floquet_integrate (src/integrator.cpp:869-889) of spec (normalised by
 * spec->dt / spec->n_samples) with opt's yi, t0, tf and w. */
ST_API int st_floquet_integrate(const st_floquet* fd, const st_spectrum* spec, const st_options* opt,
                                double b_end[3]);

/* Waits for all work queued on the handle's stream. */
ST_API int st_synchronize(st_handle* h);

/* Library identification: "spintrack gfx950 <variant list>". */
ST_API const char* st_version(void);

#ifdef __cplusplus
}
#endif
#endif
