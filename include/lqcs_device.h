/*
 * lqcs_device.h -- thin C-ABI over the MI355X (gfx950) statevector engine.
 *
 * The C++20 host library (include/lqcs/...) and any FFI (ctypes, cgo, JNI)
 * reach the HIP kernels only through these entry points. Plain pointers and
 * sizes; no HIP or torch types. Every call returns an lqcs_status; the text
 * of the last error on the calling thread is lqcs_dev_last_error().
 *
 * State layout in HBM: 2^n complex128 amplitudes, interleaved (re, im), index
 * bit q = qubit q (little endian), 256-byte aligned -- the device image of
 * the reference's aligned_vector<complex_t> (include/lqcs/core/types.hpp:11-16).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * LiteQuantumComputingSimulator).
 */
#ifndef LQCS_DEVICE_H
#define LQCS_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int lqcs_status;
enum {
    LQCS_OK = 0,
    LQCS_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    LQCS_ERR_OUT_OF_RANGE = 2,     /* std::out_of_range */
    LQCS_ERR_LOGIC = 3,            /* std::logic_error */
    LQCS_ERR_OUT_OF_MEMORY = 4,    /* device allocation failed */
    LQCS_ERR_DEVICE = 5,           /* HIP runtime error */
    LQCS_ERR_NO_DEVICE = 6         /* no usable gfx950 device */
};

/* Largest register the device engine accepts (2^34 amplitudes = 256 GiB). */
#define LQCS_DEVICE_MAX_QUBITS 34

typedef struct lqcs_dev_state lqcs_dev_state;

/* ---- context ------------------------------------------------------------ */
const char* lqcs_dev_last_error(void);
/* Selects the device and creates the engine stream. Idempotent. */
lqcs_status lqcs_dev_init(int device);
/* Device name / CU count / HBM bytes of the active device. */
lqcs_status lqcs_dev_info(char* name, size_t name_len, int* cus, uint64_t* hbm_bytes);
lqcs_status lqcs_dev_synchronize(void);

/* ---- state: replaces Statevector (include/lqcs/backend/statevector.hpp:12-37)
 * lqcs_state_create = Statevector(n) (src/backend/statevector.cpp:11-25). */
lqcs_status lqcs_state_create(uint32_t n_qubits, lqcs_dev_state** out);
lqcs_status lqcs_state_destroy(lqcs_dev_state* s);
lqcs_status lqcs_state_num_qubits(const lqcs_dev_state* s, uint32_t* n_qubits);
/* Reinitialise to the basis state |index>. */
lqcs_status lqcs_state_set_basis(lqcs_dev_state* s, uint64_t index);
/* Host <-> device copies of the whole register, (re, im) interleaved. */
lqcs_status lqcs_state_upload(lqcs_dev_state* s, const double* amps);
lqcs_status lqcs_state_download(const lqcs_dev_state* s, double* amps);
lqcs_status lqcs_state_copy(lqcs_dev_state* dst, const lqcs_dev_state* src);
/* Raw device address of the amplitudes (for zero-copy interop). */
lqcs_status lqcs_state_device_ptr(lqcs_dev_state* s, void** ptr);

/* ---- operation stream: replaces the kernel seam lqcs::kernels::*
 * (src/backend/kernels/kernels.hpp:13-62), fed by the dispatcher
 * apply_instruction (src/backend/statevector_simulator.cpp:22-85). */
enum {
    LQCS_OP_SINGLE = 0,           /* apply_single_qubit: targets[0], 2x2 */
    LQCS_OP_DIAGONAL_SINGLE = 1,  /* apply_diagonal_single_qubit: targets[0], d0,d1 */
    LQCS_OP_CONTROLLED_SINGLE = 2,/* apply_controlled_single_qubit: 1 control, 2x2 */
    LQCS_OP_SWAP = 3,             /* apply_swap: targets[0], targets[1] */
    LQCS_OP_TWO_QUBIT = 4,        /* apply_two_qubit: 4x4, local bit0 <-> targets[0] */
    LQCS_OP_CONTROLLED_UNITARY = 5,    /* apply_controlled_unitary: c controls, k targets */
    LQCS_OP_CONTROLLED_PERMUTATION = 6,/* apply_controlled_permutation */
    LQCS_OP_PAULI_ROTATION = 7    /* apply_pauli_rotation: x_mask, zy_mask, n_y, theta */
};

typedef struct lqcs_op {
    uint32_t kind;
    uint32_t n_controls;
    uint32_t n_targets;
    uint32_t qubit_offset;  /* into lqcs_program.qubits: controls, then targets */
    uint64_t data_offset;   /* complex entries (matrix / d0,d1) or perm entries */
    uint64_t x_mask;        /* PAULI_ROTATION: global X/Y bits */
    uint64_t zy_mask;       /* PAULI_ROTATION: global Z/Y bits */
    uint32_t n_y;
    uint32_t reserved;
    double theta;
} lqcs_op;

typedef struct lqcs_program {
    const lqcs_op* ops;
    uint64_t n_ops;
    const uint32_t* qubits;
    uint64_t n_qubits;
    const double* cplx;     /* (re, im) pairs */
    uint64_t n_cplx;
    const uint64_t* perm;
    uint64_t n_perm;
} lqcs_program;

/* Execution plan choice for lqcs_state_apply. */
enum {
    LQCS_PLAN_GATE_BY_GATE = 0,  /* one reference-seam kernel per op */
    LQCS_PLAN_GREEDY = 1,        /* tiled passes, first-fit grouping in program order */
    LQCS_PLAN_LEARNED = 2,       /* the installed cost model's pick among GREEDY and GAIN
                                    plans with 2..4 low qubits (specialized passes: 3..5 low
                                    qubits on 12- and 13-qubit tiles) and, from 12 qubits, DENSE
                                    blocks of 4..6 qubits; GREEDY when none is installed */
    LQCS_PLAN_AUTO = 3,          /* engine default (GREEDY) */
    LQCS_PLAN_GAIN = 4,          /* tiled passes, tile sets grown by marginal gain */
    LQCS_PLAN_DENSE = 5          /* ops grouped into blocks of tile_qubits qubits (1..6; default
                                    5, or the register's width if smaller), each multiplied on the
                                    host into one unitary and applied in one pass (a complex GEMM
                                    on FP64 MFMA for 4..6 qubits) */
};

/* Kernel for tiled passes, per call (lqcs_exec_options.tile_kernel). */
enum {
    LQCS_TILE_DEFAULT = 0,   /* the one lqcs_set_tile_kernel selected */
    LQCS_TILE_LDS = 1,
    LQCS_TILE_REGISTER = 2,
    LQCS_TILE_STREAM = 3,    /* register tile, next tile prefetched into registers */
    LQCS_TILE_LAND = 4,      /* register tile, next tile prefetched into LDS by DMA */
    LQCS_TILE_SPECIALIZED = 5 /* register tile, one kernel per pass structure compiled at run time
                                 (hipRTC) and cached for the process; 12-qubit tiles, other
                                 passes run on the register tile */
};

typedef struct lqcs_exec_options {
    int32_t plan;         /* LQCS_PLAN_* */
    int32_t tile_qubits;  /* tile width L (0 = engine default) */
    int32_t low_qubits;   /* contiguous low qubits kept in every tile (0 = default) */
    int32_t tile_kernel;  /* LQCS_TILE_* */
} lqcs_exec_options;

/* Applies the whole op stream in order, in place. opts may be NULL. */
lqcs_status lqcs_state_apply(lqcs_dev_state* s, const lqcs_program* prog,
                             const lqcs_exec_options* opts);
/* ---- captured launch sequences (no counterpart in the reference): the
 * launches lqcs_state_apply would issue for (s, prog, opts), recorded once
 * into a HIP graph with the tile programs in device memory owned by the
 * graph, then replayed by lqcs_graph_launch (asynchronous; ordered with
 * the engine's other work; lqcs_dev_synchronize waits). The graph's kernels
 * address the register s held at capture: lqcs_graph_launch takes the
 * register to replay on and returns LOGIC unless it holds that same device
 * buffer at the same width (a register replaced since, e.g. by assignment,
 * is refused rather than written through a stale address). The tables of
 * stand-alone controlled-unitary and controlled-permutation steps (dense
 * blocks included) live in device memory the graph owns too. */
typedef struct lqcs_graph lqcs_graph;
lqcs_status lqcs_graph_create(lqcs_dev_state* s, const lqcs_program* prog, const lqcs_exec_options* opts,
                              lqcs_graph** out);
lqcs_status lqcs_graph_launch(lqcs_graph* g, lqcs_dev_state* s);
lqcs_status lqcs_graph_destroy(lqcs_graph* g);

/* Plan summary of the last lqcs_state_apply on this thread: HBM passes and
 * kernel launches it issued. */
lqcs_status lqcs_last_plan_stats(uint64_t* passes, uint64_t* launches);

/* Plan of lqcs_state_apply without running it (no device needed): number of
 * steps (HBM passes), ops per step, each step's qubit mask (a tile's qubits, a
 * dense block's qubits, 0 for a stand-alone gate-by-gate step) and its
 * register-tile phase count (0 when the step runs on another kernel); arrays
 * hold max_steps entries. */
lqcs_status lqcs_plan_describe(uint32_t n_qubits, const lqcs_program* prog, const lqcs_exec_options* opts,
                               uint64_t* n_steps, uint64_t* ops_per_step, uint64_t* tile_masks,
                               uint64_t* phases_per_step, uint64_t max_steps);

/* Specialized kernel of one planned step of lqcs_state_apply (no device
 * needed; no counterpart in the reference): the kernel text the engine
 * compiles for it (after the device templates of spec_pass.hpp) and the op
 * words it reads, as (re, im) pairs of doubles. text_len is 0 when the step
 * runs on another kernel. Null text or words (or too small a words buffer)
 * skip the copy; text receives at most text_max - 1 characters and a NUL.
 * opts selects the plan as for lqcs_state_apply. */
lqcs_status lqcs_plan_specialized_pass(uint32_t n_qubits, const lqcs_program* prog, const lqcs_exec_options* opts,
                                       uint64_t step, char* text, size_t text_max, size_t* text_len, double* words,
                                       uint64_t words_max, uint64_t* n_words);

/* Number of specialized pass kernels compiled and loaded in this process
 * (cached by kernel text; no counterpart in the reference). */
lqcs_status lqcs_specialized_kernels(uint64_t* cached);

/* Kernel used for tiled passes: 1 = register-resident tile, one workgroup
 * per tile (default); 2 = register tile, streaming (each workgroup walks
 * tiles and prefetches the next one into registers); 3 = register tile,
 * landing (one workgroup per CU, the next tile prefetched into LDS by DMA);
 * 4 = register tile specialized per pass structure (LQCS_TILE_SPECIALIZED);
 * 0 = LDS tile. A switch for comparisons; results agree to rounding. */
lqcs_status lqcs_set_tile_kernel(int kernel);

/* Dense controlled unitaries (reference src/backend/kernels/gate_kernels.cpp:143-167)
 * with k >= min_targets targets, k <= 6, run as complex GEMMs on FP64 MFMA;
 * the others on the vector kernels. Default 4; 7 turns MFMA off. A switch
 * for comparisons; results agree to rounding. */
lqcs_status lqcs_set_dense_mfma_min_targets(uint32_t min_targets);

/* Launch timing (no counterpart in the reference): while enabled, every
 * tiled pass and every gate-by-gate kernel of lqcs_state_apply (and of the
 * batched trajectory and density-matrix runs built on it) is bracketed by
 * HIP events on the engine stream. lqcs_profile_read synchronises and
 * returns launch counts, summed kernel milliseconds and the amplitudes the
 * tiled passes passed over (registers x 2^n per pass; each pass reads and
 * writes 16 bytes per amplitude), then clears them. Any pointer may be NULL. */
lqcs_status lqcs_profile_enable(int enable);
lqcs_status lqcs_profile_read(uint64_t* tile_launches, double* tile_ms, uint64_t* other_launches,
                              double* other_ms, uint64_t* tile_amps);

/* Launch times of the same recording, one entry per launch in order
 * (tiled[i] = 1 for a tiled pass), then cleared; n_launches receives the
 * count, arrays hold max_launches entries. */
lqcs_status lqcs_profile_read_launches(double* ms, uint8_t* tiled, uint64_t max_launches,
                                       uint64_t* n_launches);

/* ---- learned cost model (no counterpart in the reference; it steers the
 * planner in place of the reference's fixed greedy fusion). A multilayer
 * perceptron F -> 64 -> 64 -> 1 (ReLU), FP64, trained and evaluated on the
 * device. Parameters are packed as W1 [F][64] (input-major), b1 [64],
 * W2 [64][64], b2 [64], w3 [64], b3. */
#define LQCS_COST_FEATURES 24
lqcs_status lqcs_mlp_num_params(uint32_t n_features, uint64_t* n_params);
/* y[r] for each of n_rows rows of x (row-major, n_features columns). */
lqcs_status lqcs_mlp_forward(uint32_t n_features, const double* params, const double* x,
                             uint64_t n_rows, double* y);
/* Mean squared error over the rows and its gradient (packed like params). */
lqcs_status lqcs_mlp_loss_grad(uint32_t n_features, const double* params, const double* x,
                               const double* t, uint64_t n_rows, double* loss, double* grad);
/* Adam on shuffled minibatches (mt19937_64(seed) permutation per epoch),
 * params updated in place; final_loss = loss over all rows afterwards. */
lqcs_status lqcs_mlp_train(uint32_t n_features, double* params, const double* x, const double* t,
                           uint64_t n_rows, uint32_t epochs, uint32_t batch_rows, double lr,
                           uint64_t seed, double* final_loss);
/* Installs the model LQCS_PLAN_LEARNED consults: predicted milliseconds of a
 * plan step = exp(t_mean + t_scale * net((features - x_mean) / x_scale)).
 * params NULL uninstalls it. */
lqcs_status lqcs_costmodel_install(uint32_t n_features, const double* params, const double* x_mean,
                                   const double* x_scale, double t_mean, double t_scale);
/* The same for the model of one tile kernel: LQCS_TILE_REGISTER (the model
 * above) or LQCS_TILE_SPECIALIZED, which LQCS_PLAN_LEARNED consults for the
 * calls that select specialized passes; its candidates then include 13-qubit
 * tiles and five low qubits. */
lqcs_status lqcs_costmodel_install_kernel(int32_t tile_kernel, uint32_t n_features, const double* params,
                                          const double* x_mean, const double* x_scale, double t_mean,
                                          double t_scale);
/* Features of every step of the plan lqcs_state_apply would run for these
 * options on a register of n_qubits (batch members per launch), in step
 * order, LQCS_COST_FEATURES values per step; arrays hold max_steps rows. */
lqcs_status lqcs_plan_features(uint32_t n_qubits, const lqcs_program* prog, const lqcs_exec_options* opts,
                               uint32_t batch, double* features, uint64_t max_steps, uint64_t* n_steps);
/* Plan kind and low-qubit count the last apply on this thread ran, and the
 * cost model's predicted milliseconds for it (-1 when no model decided). */
lqcs_status lqcs_last_plan_choice(int32_t* plan, int32_t* low_qubits, double* predicted_ms);
/* Tile width of that plan (0: gate by gate). */
lqcs_status lqcs_last_plan_tile_qubits(int32_t* tile_qubits);

/* ---- reductions and measurement ---------------------------------------- */
/* Statevector::norm (statevector.cpp:27-31) squared, before the sqrt. */
lqcs_status lqcs_state_norm_sq(const lqcs_dev_state* s, double* out);
/* max_i |a_i - b_i| of two registers of one width, on the device (no
 * counterpart in the reference: compares results where a host copy of the
 * register is impractical; NaN when any difference is NaN). */
lqcs_status lqcs_state_max_abs_diff(const lqcs_dev_state* a, const lqcs_dev_state* b, double* out);
/* Statevector::probabilities (statevector.cpp:33-37) into host memory. */
lqcs_status lqcs_state_probabilities(const lqcs_dev_state* s, double* out);
/* Copies amplitudes [first, first + count) to out (interleaved re, im):
 * the device side of Statevector::operator[] and DensityMatrix::element. */
lqcs_status lqcs_state_read(const lqcs_dev_state* s, uint64_t first, uint64_t count, double* out);

/* Density matrices (reference include/lqcs/backend/density_matrix.hpp:17-44,
 * src/backend/density_matrix.cpp:19-124): an n-qubit rho is a 2n-qubit register,
 * rho[r][c] at index r | (c << n) (ket bits low, bra bits high), evolved by
 * lqcs_state_apply with U on the ket qubits, conj(U) on the bra qubits and
 * each Kraus channel as its 4x4 superoperator sum_i K_i (x) conj(K_i) on the
 * qubit pair (q, q + n). */
/* out[r] = Re rho[r][r], 2^n doubles (DensityMatrix::probabilities, density_matrix.cpp:100-105). */
lqcs_status lqcs_dm_diagonal(const lqcs_dev_state* rho, double* out);
/* <psi|rho|psi> for an n-qubit register psi (DensityMatrix::fidelity, density_matrix.cpp:107-124). */
lqcs_status lqcs_dm_fidelity(const lqcs_dev_state* rho, const lqcs_dev_state* psi, double* out);
/* Statevector::expectation_value (statevector.cpp:39-78) with the string
 * already turned into global masks (leftmost char = qubit n-1). */
lqcs_status lqcs_state_expectation_pauli(const lqcs_dev_state* s, uint64_t x_mask,
                                         uint64_t zy_mask, uint32_t n_y, double* out);
/* algorithms::expectation (src/algorithms/vqe.cpp:11-17): sum of coeff_t *
 * <P_t> over n_terms Pauli terms in one call. */
lqcs_status lqcs_state_expectation_terms(const lqcs_dev_state* s, const uint64_t* x_masks,
                                         const uint64_t* zy_masks, const uint32_t* n_ys,
                                         const double* coeffs, uint64_t n_terms, double* out);
/* No counterpart in the reference: lqcs_state_expectation_terms for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments and result;
 * it calls no hipMalloc/hipFree: the staged terms and partial sums live in one
 * static device array of about 8 MiB per GPU, held until the process ends.
 * At most LQCS_NOALLOC_MAX_TERMS terms (more: INVALID_ARGUMENT). Every call
 * shares that array; calls hold the engine lock and return after their work
 * completes, so calls through this library never overlap on one GPU (work
 * that did overlap would race on the array). */
#define LQCS_NOALLOC_MAX_TERMS 512
lqcs_status lqcs_state_expectation_terms_noalloc(const lqcs_dev_state* s, const uint64_t* x_masks,
                                                 const uint64_t* zy_masks, const uint32_t* n_ys,
                                                 const double* coeffs, uint64_t n_terms, double* out);
/* run_sampling (statevector_simulator.cpp:88-119): for each uniform draw u,
 * the index upper_bound(cumulative, u * total). */
lqcs_status lqcs_state_sample(const lqcs_dev_state* s, const double* draws, uint64_t n_draws,
                              uint64_t* out_indices);
/* collapse_qubit / reset_qubit (gate_kernels.cpp:227-260). */
lqcs_status lqcs_state_collapse(lqcs_dev_state* s, uint32_t qubit, double u, int* outcome);
lqcs_status lqcs_state_reset_qubit(lqcs_dev_state* s, uint32_t qubit, double u);
/* kraus_probability (gate_kernels.cpp:262-277); k = 4 complex (8 doubles). */
lqcs_status lqcs_state_kraus_probability(const lqcs_dev_state* s, uint32_t qubit,
                                         const double* k, double* out);

/* ---- per-shot trajectories: replaces run_per_shot + sample_kraus_branch +
 * apply_trajectory_noise (statevector_simulator.cpp:122-190). The op stream
 * carries one unitary op per circuit instruction plus marker ops below; draws
 * hold n_shots * draws_per_shot uniforms in the reference's consumption order
 * (shot s uses [s*D, (s+1)*D)). Several shots run as one batch on device. */
enum {
    LQCS_TOP_UNITARY = 0,  /* apply ops[op_index]; then noise on its qubits */
    LQCS_TOP_MEASURE = 1,  /* collapse qubit, write clbit */
    LQCS_TOP_RESET = 2     /* reset qubit */
};
typedef struct lqcs_traj_step {
    uint32_t kind;
    uint32_t qubit;        /* MEASURE / RESET */
    uint32_t clbit;        /* MEASURE */
    uint32_t op_index;     /* UNITARY: index into prog->ops */
    uint32_t noise_qubit_offset;  /* UNITARY: qubits the channels act on */
    uint32_t n_noise_qubits;
} lqcs_traj_step;

typedef struct lqcs_channel {
    uint32_t n_ops;
    const double* ops;     /* n_ops * 4 complex, row-major 2x2, (re,im) */
} lqcs_channel;

/* Trajectory kernels (the `kernel` argument). PERSISTENT runs every shot in
 * one launch, a workgroup per shot with its register in LDS, all branch
 * decisions on the device; it needs n_qubits <= 12 and ops of at most two
 * targets (else INVALID_ARGUMENT). SHARED_RUNS keeps max_batch registers in
 * HBM and runs the ops between state-dependent events as tiled passes over
 * the batch. AUTO takes PERSISTENT where it applies. */
enum {
    LQCS_TRAJ_AUTO = 0,
    LQCS_TRAJ_SHARED_RUNS = 1,
    LQCS_TRAJ_PERSISTENT = 2
};

lqcs_status lqcs_trajectories_run(uint32_t n_qubits, uint32_t n_clbits,
                                  const lqcs_program* prog, const lqcs_traj_step* steps,
                                  uint64_t n_steps, const uint32_t* noise_qubits,
                                  const lqcs_channel* channels, uint32_t n_channels,
                                  const double* draws, uint64_t draws_per_shot,
                                  uint64_t n_shots, uint32_t max_batch, uint32_t kernel,
                                  uint8_t* clbits_out);

/* ---- adjoint differentiation of <H> for Pauli-rotation parameters
 * (replaces algorithms::parameter_shift_gradient, src/algorithms/vqe.cpp:68-94,
 * for circuits whose parameters are RX/RY/RZ/PauliRotation angles).
 * prog holds the bound circuit; param_op[p] is the op index of parameter p,
 * whose generator is its PAULI_ROTATION masks (RX/RY/RZ lowered to 1-qubit
 * Pauli rotations by the caller). Writes <H> to energy and dE/dtheta_p. */
lqcs_status lqcs_adjoint_gradient(uint32_t n_qubits, const lqcs_program* prog,
                                  const uint64_t* param_op, uint64_t n_params,
                                  const uint64_t* x_masks, const uint64_t* zy_masks,
                                  const uint32_t* n_ys, const double* coeffs,
                                  uint64_t n_terms, double* energy, double* grad);

/* Batched adjoint over many circuits of possibly different widths in shared
 * launches (one launch per op step across the batch). Circuit c uses
 * progs[c], n_qubits[c], params param_op[param_off[c] .. param_off[c+1]) and
 * terms [term_off[c], term_off[c+1]); grads are written at param_off[c]. */
lqcs_status lqcs_adjoint_gradient_batched(uint32_t n_circuits, const uint32_t* n_qubits,
                                          const lqcs_program* progs,
                                          const uint64_t* param_op, const uint64_t* param_off,
                                          const uint64_t* x_masks, const uint64_t* zy_masks,
                                          const uint32_t* n_ys, const double* coeffs,
                                          const uint64_t* term_off, double* energies,
                                          double* grads);

/* Captured adjoint plan: the batch of lqcs_adjoint_gradient_batched (same
 * argument layout) with its structure fixed and its angles free. Create
 * builds and places every launch group once and captures the whole launch
 * sequence in one HIP graph; each run writes theta[p] (indexed like
 * param_op, n_theta = param_off[n_circuits]) into parameter p's rotation,
 * replays the graph and writes energies and grads as the batched call does.
 * The angles the programs hold for parameter ops are replaced at every run;
 * other ops keep theirs. Ops of kind CONTROLLED_UNITARY or
 * CONTROLLED_PERMUTATION are refused (INVALID_ARGUMENT). The plan holds its
 * registers (the largest launch group's) until destroyed.
 * No counterpart in the reference: it serves optimiser loops that evaluate
 * the same ansatz batch at many angles (reference loop: vqe_gradient_descent,
 * src/algorithms/vqe.cpp:96-133). */
typedef struct lqcs_adjoint_plan lqcs_adjoint_plan;
lqcs_status lqcs_adjoint_plan_create(uint32_t n_circuits, const uint32_t* n_qubits,
                                     const lqcs_program* progs, const uint64_t* param_op,
                                     const uint64_t* param_off, const uint64_t* x_masks,
                                     const uint64_t* zy_masks, const uint32_t* n_ys,
                                     const double* coeffs, const uint64_t* term_off,
                                     lqcs_adjoint_plan** out);
lqcs_status lqcs_adjoint_plan_run(lqcs_adjoint_plan* plan, const double* theta, uint64_t n_theta,
                                  double* energies, double* grads);
lqcs_status lqcs_adjoint_plan_destroy(lqcs_adjoint_plan* plan);

#ifdef __cplusplus
}
#endif
#endif
