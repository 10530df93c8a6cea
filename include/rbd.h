/*
 * rbd.h -- C ABI of the MI355X batched rigid-body-dynamics engine (libpinocchio_amd.so).
 *
 * The engine evaluates the reference's dynamics algorithms over a batch of B
 * configurations of one model in a single call, on one GPU.  Each batched entry point
 * replaces a loop of the reference's single-configuration algorithm over the columns of
 * Eigen matrices, i.e. what pinocchio's rneaInParallel / abaInParallel do with OpenMP:
 *
 *   rbd_rnea              rneaInParallel            algorithm/parallel/rnea.hpp:39-83
 *                         (per column: rnea         algorithm/rnea.hxx:117-161, 773)
 *   rbd_aba               abaInParallel             algorithm/parallel/aba.hpp:40-84
 *                         (per column: aba WORLD    algorithm/aba.hxx:242-293, 926-948)
 *   rbd_nonlinear_effects nonLinearEffects          algorithm/rnea.hpp:105, rnea.hxx:302
 *   rbd_generalized_gravity computeGeneralizedGravity algorithm/rnea.hpp:133, rnea.hxx:412
 *   rbd_static_torque     computeStaticTorque       algorithm/rnea.hpp:163, rnea.hxx:454-490
 *   rbd_rnea_fext, rbd_aba_fext
 *                         rnea / aba with fext      algorithm/rnea.hpp:60-79, aba.hpp:63-85
 *   rbd_crba              crba                      algorithm/crba.hxx:498-548, 602-617
 *   rbd_cholesky_decompose cholesky::decompose      algorithm/cholesky.hxx:17-62
 *   rbd_cholesky_solve    cholesky::solve           algorithm/cholesky.hxx:466-520
 *   rbd_cholesky_apply    cholesky::Uv/Utv/Uiv/Utiv/Mv algorithm/cholesky.hxx:64-465
 *   rbd_rnea_derivatives  computeRNEADerivatives    algorithm/rnea-derivatives.hxx:472-541, 756-766
 *   rbd_aba_derivatives   computeABADerivatives     algorithm/aba-derivatives.hxx:380-453, 900-923
 *   rbd_compute_minverse  computeMinverse           algorithm/aba.hxx:608-902
 *   rbd_forward_kinematics forwardKinematics        algorithm/kinematics.hxx:121-145, 209-238
 *   rbd_joint_jacobians   computeJointJacobians     algorithm/jacobian.hxx:23-62
 *   rbd_joint_jacobians_time_variation
 *                         computeJointJacobiansTimeVariation  algorithm/jacobian.hxx:673-740
 *   rbd_frames            updateFramePlacements,    algorithm/frames.hxx:15-35,
 *                         getFrameJacobian          algorithm/frames.hxx:156-175
 *   rbd_frame_jacobians_time_variation
 *                         getFrameJacobianTimeVariation  algorithm/frames.hxx:293-368
 *   rbd_frame_motions     getFrameVelocity,         algorithm/frames.hxx:71-99,
 *                         getFrameAcceleration,     algorithm/frames.hxx:102-130,
 *                         getFrameClassicalAcceleration  algorithm/frames.hxx:133-149
 *   rbd_constraint_dynamics constraintDynamics      algorithm/constrained-dynamics.hxx:193-500
 *   rbd_delassus          computeDelassusMatrix     algorithm/delassus.hpp:41-60, delassus.hxx:118-300
 *   rbd_damped_delassus_inverse
 *                         computeDampedDelassusMatrixInverse
 *                                                   algorithm/delassus.hpp:84-102, delassus.hxx:800-836
 *   rbd_contact_cholesky_solve
 *                         ContactCholeskyDecomposition compute + solve
 *                                                   algorithm/contact-cholesky.hpp:292-400
 *   rbd_pgs_solve         PGSContactSolverTpl::solve algorithm/pgs-solver.hpp:16-59, pgs-solver.hxx:19-141
 *   rbd_admm_solve        ADMMContactSolverTpl::solve algorithm/admm-solver.hpp:120-316, admm-solver.hxx:22-335
 *   rbd_impulse_dynamics  impulseDynamics(q, v_before, J, r_coeff, inv_damping)
 *                                                   algorithm/contact-dynamics.hxx:264-331
 *   rbd_constrained_forward_dynamics
 *                         forwardDynamics(q, v, tau, J, gamma, inv_damping)
 *                                                   algorithm/contact-dynamics.hxx:90-167
 *   rbd_multi_rnea_aba    (no counterpart: one launch over instances of several models)
 *   rbd_mlp_*             (no counterpart: learned forward-dynamics surrogate, checked by
 *                          the RNEA residual with exact ABA as fallback)
 *
 * Conventions (identical to the reference): 6-vectors are [linear; angular]; free-flyer
 * configuration is [x, y, z, qx, qy, qz, qw]; joint 0 is the universe; parents[i] < i.
 *
 * Array layout.  A batched vector argument (q: nq x B, v/a/tau/ddq: nv x B) is a matrix
 * with one configuration per column.  By default it is column-major like Eigen:
 * element (k, b) at p[k + b * ld] with ld >= rows.  With RBD_LAYOUT_BATCH the batch index
 * is fastest: element (k, b) at p[k * ld + b] with ld >= B (the engine's native layout).
 * A batched matrix argument (nv x nv per configuration) uses element index k = r + c * nv
 * (column-major inside the configuration) in the same two layouts.
 *
 * Pointers are device pointers (hipMalloc'd or torch tensors) unless RBD_HOST_PTRS is
 * set, in which case the engine stages through its own device buffers and synchronises.
 * Without RBD_ASYNC every call synchronises its stream before returning, like the
 * reference's synchronous calls; with RBD_ASYNC it returns once the work is enqueued.
 *
 * Return value: 0 on success, otherwise an RBD_E* code; rbd_last_error() then gives a
 * message (thread-local).  Shape errors mirror the reference's std::invalid_argument
 * conditions (algorithm/parallel/rnea.hpp:53-68, algorithm/parallel/aba.hpp:54-69).
 */
#ifndef RBD_H_
#define RBD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RBD_OK = 0,
  RBD_EINVAL = 1,      /* bad argument or shape (std::invalid_argument in the reference) */
  RBD_EUNSUPPORTED = 2,/* joint type / topology / dtype the engine does not handle */
  RBD_EDEVICE = 3,     /* HIP runtime error */
  RBD_ENOMEM = 4
};

enum { RBD_F32 = 0, RBD_F64 = 1 };

/* joint types */
enum { RBD_JOINT_UNIVERSE = 0, RBD_JOINT_FREEFLYER = 1, RBD_JOINT_RX = 2, RBD_JOINT_RY = 3, RBD_JOINT_RZ = 4 };

/* flags */
#define RBD_HOST_PTRS 0x1u     /* arguments are host pointers */
#define RBD_ASYNC 0x2u         /* do not synchronise before returning */
#define RBD_LAYOUT_BATCH 0x4u  /* batch index fastest (see header comment) */
#define RBD_UPPER_ONLY 0x8u    /* matrix outputs: write the upper triangle only (reference's crba) */

/* Flattened model (multibody/model.hpp:98-220).  All arrays have njoints entries, entry 0
 * being the universe.  placement[i] = R (row major, 9) then translation (3), the joint
 * placement in the parent frame; inertia[i] = mass, lever c (3), rotational inertia about
 * the centre of mass as (xx, xy, yy, xz, yz, zz).  armature has nv entries (may be NULL),
 * gravity 6 (may be NULL for (0, 0, -9.81, 0, 0, 0)).  The engine copies what it needs. */
typedef struct rbd_model_desc {
  int32_t njoints;
  const int32_t* parents;
  const int32_t* jtype;
  const double* placement;
  const double* inertia;
  const double* armature;
  const double* gravity;
} rbd_model_desc;

typedef struct rbd_ctx rbd_ctx;
typedef struct rbd_model rbd_model;

const char* rbd_last_error(void);
const char* rbd_version(void);

/* Context: owns a HIP stream on `device` and the staging buffers.  Replaces the
 * ModelPool + OpenMP thread team of the reference (multibody/pool/model.hpp:19-165,
 * algorithm/parallel/omp.hpp:12-16); no global state is mutated.  The stream is a
 * blocking stream: its work is ordered with work on the legacy default stream (HIP's
 * null stream, PyTorch's default stream); callers that produce inputs or consume outputs
 * on other streams order them with events on rbd_ctx_stream(). */
int rbd_ctx_create(int device, rbd_ctx** out);
int rbd_ctx_destroy(rbd_ctx* ctx);
/* The stream the context enqueues on (a hipStream_t). */
void* rbd_ctx_stream(rbd_ctx* ctx);
int rbd_ctx_synchronize(rbd_ctx* ctx);

/* Model upload (ModelTpl::addJoint, multibody/model.hxx:61-170, flattened).  Any tree of
 * revolute X / Y / Z joints with free-flyer joints at the roots is accepted: trees the
 * library has prebuilt kernels for (compiled for their exact topology) run on those, every
 * other tree on the run-time-topology kernels, which read the tree as data.  Fails with
 * RBD_EUNSUPPORTED for other joint types or a free flyer below another joint. */
int rbd_model_create(rbd_ctx* ctx, const rbd_model_desc* desc, rbd_model** out);
/* As rbd_model_create; with RBD_MODEL_RUNTIME_TOPOLOGY the model runs on the
 * run-time-topology kernels even when its tree is prebuilt (for comparisons).  No
 * reference counterpart. */
#define RBD_MODEL_RUNTIME_TOPOLOGY 0x1u
int rbd_model_create_ex(rbd_ctx* ctx, const rbd_model_desc* desc, uint32_t model_flags, rbd_model** out);
/* *prebuilt = 1 when the model's FP32 sweeps run on kernels compiled for its topology, 0
 * when on the run-time-topology kernels.  (FP64 ABA and FP64 derivatives always use the
 * run-time-topology kernels.)  No reference counterpart. */
int rbd_model_kernels(const rbd_model* model, int32_t* prebuilt);
int rbd_model_destroy(rbd_model* model);
/* nq, nv, njoints and whether the model is DFS-compact (crba.hxx:573-595). */
int rbd_model_dims(const rbd_model* model, int32_t* nq, int32_t* nv, int32_t* njoints, int32_t* compact);

/* tau = rnea(q, v, a) per configuration.  Replaces rneaInParallel. */
int rbd_rnea(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
             const void* v, int64_t ldv, const void* a, int64_t lda, void* tau, int64_t ldtau, uint32_t flags);
/* nonLinearEffects (rnea.hpp:105, rnea.hxx:302): C(q, v) v + g(q) = rnea(q, v, 0), and
 * computeGeneralizedGravity (rnea.hpp:133, rnea.hxx:412): g(q) = rnea(q, 0, 0);
 * the zero inputs are not read from the caller.  Layouts and flags as rbd_rnea. */
int rbd_nonlinear_effects(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                          const void* v, int64_t ldv, void* nle, int64_t ldnle, uint32_t flags);
int rbd_generalized_gravity(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                            int64_t ldq, void* g, int64_t ldg, uint32_t flags);
/* computeStaticTorque (rnea.hpp:163, rnea.hxx:454-490): g(q) with the external forces,
 * rnea(q, 0, 0, fext); fext as rbd_rnea_fext (6 (njoints - 1) rows per configuration,
 * [linear; angular] in each joint's frame).  Run-time-topology kernels for every tree. */
int rbd_static_torque(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                      const void* fext, int64_t ldfext, void* tau, int64_t ldtau, uint32_t flags);

/* ddq = aba(q, v, tau) per configuration.  Replaces abaInParallel (WORLD convention; the
 * result is convention-independent up to rounding, the engine sweeps in local frames). */
int rbd_aba(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
            const void* v, int64_t ldv, const void* tau, int64_t ldtau, void* ddq, int64_t ldddq, uint32_t flags);

/* M = crba(q) per configuration: joint-space inertia matrix with armature on the
 * diagonal.  Symmetric by default; RBD_UPPER_ONLY writes the upper triangle only and
 * leaves the rest untouched (the reference's behaviour). */
int rbd_crba(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
             void* M, int64_t ldM, uint32_t flags);

/* FP64 tree-sparse factorisation M = U D U^T of each M (upper triangle read; dtype is
 * the dtype of M).  U is written as the nnz_upper strictly-upper structural non-zeros in
 * double, column by column (see rbd_model_cholesky_pattern), and D as nv doubles.  Both
 * use the layout given by flags with element index = position in the packed list. */
int rbd_cholesky_decompose(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* M,
                           int64_t ldM, double* U, int64_t ldU, double* D, int64_t ldD, uint32_t flags);
/* Number of strictly-upper structural non-zeros and their (row, col) list, in packed order. */
int rbd_model_cholesky_pattern(const rbd_model* model, int32_t* nnz, int32_t* rows, int32_t* cols);
/* x <- M^{-1} x with a factor from rbd_cholesky_decompose (x has dtype `dtype`). */
int rbd_cholesky_solve(rbd_ctx* ctx, const rbd_model* model, int64_t batch, const double* U, int64_t ldU,
                       const double* D, int64_t ldD, int dtype, void* x, int64_t ldx, uint32_t flags);

/* rnea(model, data, q, v, a, fext) and aba(model, data, q, v, tau, fext) (rnea.hpp:60-79,
 * aba.hpp:63-85) per configuration: fext holds 6 (njoints - 1) rows per configuration, the
 * external force on each joint 1.. as [linear; angular] in that joint's frame.  Run on the
 * run-time-topology kernels for every tree (the prebuilt kernels take no forces). */
int rbd_rnea_fext(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                  const void* v, int64_t ldv, const void* a, int64_t lda, const void* fext, int64_t ldfext, void* tau,
                  int64_t ldtau, uint32_t flags);
int rbd_aba_fext(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                 const void* v, int64_t ldv, const void* tau, int64_t ldtau, const void* fext, int64_t ldfext,
                 void* ddq, int64_t ldddq, uint32_t flags);

/* The factor's products and triangular solves, in place on x (nv rows, FP64), with U and D
 * from rbd_cholesky_decompose: cholesky::Uv (x <- U x), Utv (U^T x), Uiv (U^-1 x), Utiv
 * (U^-T x) and Mv (M x = U D U^T x) (cholesky.hpp:86-150, cholesky.hxx:64-465). */
enum { RBD_CHOL_UV = 0, RBD_CHOL_UTV = 1, RBD_CHOL_UIV = 2, RBD_CHOL_UTIV = 3, RBD_CHOL_MV = 4 };
int rbd_cholesky_apply(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int op, const double* U, int64_t ldU,
                       const double* D, int64_t ldD, double* x, int64_t ldx, uint32_t flags);

/* computeRNEADerivatives: dtau_dq, dtau_dv, dtau_da (= M) per configuration; any output
 * may be NULL.  Structurally zero blocks are written as zeros.  Requires gravity with a
 * zero angular part, like the reference (rnea-derivatives.hxx:494-496). */
int rbd_rnea_derivatives(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                         int64_t ldq, const void* v, int64_t ldv, const void* a, int64_t lda, void* dtau_dq,
                         void* dtau_dv, void* dtau_da, int64_t ldout, uint32_t flags);

/* computeABADerivatives: ddq_dq, ddq_dv, ddq_dtau (= M^{-1}) per configuration, and
 * optionally ddq (may be NULL).  FP32 or FP64 (every step in the given precision). */
int rbd_aba_derivatives(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                        int64_t ldq, const void* v, int64_t ldv, const void* tau, int64_t ldtau, void* ddq_dq,
                        void* ddq_dv, void* ddq_dtau, int64_t ldout, void* ddq, int64_t ldddq, uint32_t flags);

/* rbd_aba_derivatives without allocating: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight control or
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * rbd_aba_derivatives; FP32, device pointers, models with prebuilt kernels.  The factor of
 * M and ddq go to a file-scope device array of 64 MiB that each GPU holds until the
 * process ends (its address is taken once per device when a context is created), so the
 * batch is limited to 64 MiB / (4 (nnz_upper + 2 nv)) configurations (44,979 for the
 * 36-DoF humanoid; larger batches fail with RBD_EINVAL).  Every call on a GPU shares that
 * array: calls that overlap on one GPU (several contexts or threads) race. */
int rbd_aba_derivatives_noalloc(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                                int64_t ldq, const void* v, int64_t ldv, const void* tau, int64_t ldtau,
                                void* ddq_dq, void* ddq_dv, void* ddq_dtau, int64_t ldout, void* ddq,
                                int64_t ldddq, uint32_t flags);

/* forwardKinematics (kinematics.hxx:121-145, 209-238): per configuration the world
 * placement oMi of joints 1 .. njoints-1 (12 entries per joint: R row major, then the
 * translation; element index 12 (i - 1) + k), and, when v (and a) are given, each joint's
 * spatial velocity (and acceleration) in its local frame (6 entries per joint, [linear;
 * angular], element index 6 (i - 1) + k).  vout / aout / v / a may be NULL. */
int rbd_forward_kinematics(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                           const void* v, int64_t ldv, const void* a, int64_t lda, void* oMi, int64_t ldoMi,
                           void* vout, void* aout, int64_t ldva, uint32_t flags);
/* computeJointJacobians (jacobian.hxx:23-62): the 6 x nv WORLD-frame Jacobian per
 * configuration, column of dof c = oMi.act(S_c); element index r + 6 c. */
int rbd_joint_jacobians(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                        void* J, int64_t ldJ, uint32_t flags);

/* computeJointJacobiansTimeVariation (jacobian.hxx:673-740): the WORLD-frame Jacobian J
 * (optional, as rbd_joint_jacobians writes it) and its time variation dJ, 6 x nv per
 * configuration, element index r + 6 c; dJ column c of joint j = ov_j x J_c with ov_j the
 * joint's spatial velocity in WORLD. */
int rbd_joint_jacobians_time_variation(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                                       int64_t ldq, const void* v, int64_t ldv, void* J, int64_t ldJ, void* dJ,
                                       int64_t lddJ, uint32_t flags);

/* computeMinverse (aba.hxx:608-902; cholesky::computeMinv, cholesky.hxx:522-597): M(q)^-1
 * per configuration, written whole (symmetric; the reference fills the upper triangle).
 * FP32: CRBA, FP32 tree factor, unit right-hand sides; FP64: the same in double. */
int rbd_compute_minverse(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                         void* Minv, int64_t ldMinv, uint32_t flags);

/* Frames (operational points): for each frame f given by its parent joint frame_joint[f]
 * and its placement in that joint's frame (frame_placement[12 f ..]: R row major, then the
 * translation), the placement oMf = oMi[parent] * placement (updateFramePlacements,
 * frames.hxx:15-35; 12 rows per frame, element 12 f + k) and the 6 x nv frame Jacobian
 * (getFrameJacobian, frames.hxx:156-175): the support columns of the parent joint's WORLD
 * Jacobian in the requested reference frame (RBD_WORLD as they are; RBD_LOCAL_WORLD_ALIGNED
 * with the linear part taken at the frame origin; RBD_LOCAL in the frame's axes), zero
 * elsewhere (element f 6 nv + r + 6 c).  oMf or J may be NULL.  frame_joint and
 * frame_placement are host arrays (the call copies them and always synchronises). */
enum { RBD_WORLD = 0, RBD_LOCAL = 1, RBD_LOCAL_WORLD_ALIGNED = 2 };
int rbd_frames(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
               int nframes, const int32_t* frame_joint, const double* frame_placement, int reference_frame,
               void* oMf, int64_t ldoMf, void* J, int64_t ldJ, uint32_t flags);

/* getFrameJacobianTimeVariation (frames.hxx:293-368) after
 * computeJointJacobiansTimeVariation(q, v), for the same frame tables as rbd_frames: the
 * time variation of each frame's 6 x nv Jacobian in the requested reference frame (element
 * f 6 nv + r + 6 c, zero outside the parent joint's support). */
int rbd_frame_jacobians_time_variation(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                                       int64_t ldq, const void* v, int64_t ldv, int nframes, const int32_t* frame_joint,
                                       const double* frame_placement, int reference_frame, void* dJ, int64_t lddJ,
                                       uint32_t flags);

/* Frame motions after forwardKinematics(q, v[, a]) for the same frame tables as rbd_frames:
 * getFrameVelocity (frames.hxx:71-99) into vf, getFrameAcceleration (frames.hxx:102-130)
 * into af and getFrameClassicalAcceleration (frames.hxx:133-149) into acf, each 6 rows per
 * frame ([linear; angular], element 6 f + k) in the requested reference frame (RBD_LOCAL,
 * RBD_WORLD or RBD_LOCAL_WORLD_ALIGNED).  a, af and acf may be NULL (af / acf need a);
 * vf, af and acf share the leading dimension ldm.  Frame tables are host arrays (the call
 * copies them and always synchronises). */
int rbd_frame_motions(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                      const void* v, int64_t ldv, const void* a, int64_t lda, int nframes, const int32_t* frame_joint,
                      const double* frame_placement, int reference_frame, void* vf, void* af, void* acf, int64_t ldm,
                      uint32_t flags);

/* forwardDynamics with acceleration constraints J ddq + gamma = 0 (contact-dynamics.hpp:
 * 66-75, contact-dynamics.hxx:90-167): M(q) and nle(q, v) as computeAllTerms gives them,
 * then ddq = M^-1 (tau - nle + J^T lambda) with
 * lambda = (J M^-1 J^T + inv_damping I)^-1 (-J M^-1 (tau - nle) - gamma), the reference's
 * data.lambda_c.  Per configuration J is nc x nv (1 <= nc <= 12), stored column major
 * like an Eigen matrix: element (i, k) is row i + k nc of the J argument (nc nv rows per
 * configuration, leading dimension ldJ, same layout flags as the vectors); gamma (nc),
 * lambda (nc) and ddq (nv) are vectors.  FP32: CRBA and RNEA in FP32, the tree factor and
 * everything after it in FP64; FP64: all in double.  Like the reference's LLT, a
 * rank-deficient J with zero damping gives non-finite results. */
int rbd_constrained_forward_dynamics(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                                     int64_t ldq, const void* v, int64_t ldv, const void* tau, int64_t ldtau, int nc,
                                     const void* J, int64_t ldJ, const void* gamma, int64_t ldgamma,
                                     double inv_damping, void* ddq, int64_t ldddq, void* lambda, int64_t ldlambda,
                                     uint32_t flags);

/* Stream capture of engine calls (no counterpart in the reference; for control loops and
 * serving paths that replay one fixed sequence of calls): between rbd_capture_begin and
 * rbd_capture_end the calls made on `ctx` are recorded into a HIP graph instead of run;
 * rbd_graph_launch replays the graph on the context's stream, reading the recorded input
 * buffers again at every replay.  During a capture, calls must take device pointers and
 * must not need to allocate: run each call once, uncaptured, with the same sizes first
 * (that reserves the workspace).  Host arrays, frame and constraint tables, mixed batches,
 * the surrogate, model creation and destruction and rbd_ctx_synchronize are refused during
 * a capture (RBD_EINVAL).  A graph holds device addresses: destroy it before the buffers
 * and models it uses. */
typedef struct rbd_graph rbd_graph;
int rbd_capture_begin(rbd_ctx* ctx);
int rbd_capture_end(rbd_ctx* ctx, rbd_graph** graph);
int rbd_graph_launch(rbd_ctx* ctx, const rbd_graph* graph, uint32_t flags); /* RBD_ASYNC or 0 */
int rbd_graph_destroy(rbd_graph* graph);

/* Rigid constraints of constraintDynamics: RigidConstraintModel (contact-info.hpp:95-420)
 * with the RigidConstraintData field constraintDynamics reads (contact_acceleration_desired,
 * zero by default).  A CONTACT_3D constraint holds the point of joint1's frame on the point
 * of joint2's frame (3 rows), CONTACT_6D the two frames together (6 rows); joint 0 is the
 * universe (a contact with the environment).  Placements: R row major, then p. */
enum { RBD_CONTACT_3D = 0, RBD_CONTACT_6D = 1 };
typedef struct rbd_rigid_constraint {
  int32_t type;             /* RBD_CONTACT_3D or RBD_CONTACT_6D */
  int32_t joint1_id, joint2_id;
  int32_t reference_frame;  /* RBD_LOCAL or RBD_LOCAL_WORLD_ALIGNED (RBD_WORLD is refused, like the reference) */
  double joint1_placement[12], joint2_placement[12];
  double Kp[6], Kd[6];      /* Baumgarte corrector gains (BaumgarteCorrectorParameters); first 3 for 3D */
  double desired_acceleration[6]; /* contact_acceleration_desired: linear, angular */
} rbd_rigid_constraint;

/* ProximalSettingsTpl (proximal.hpp:20-60); NULL settings = the reference's defaults
 * (accuracies 1e-12, mu 0, max_iter 1). */
typedef struct rbd_proximal_settings {
  double absolute_accuracy, relative_accuracy, mu;
  int32_t max_iter;
} rbd_proximal_settings;

/* constraintDynamics(model, data, q, v, tau, contact_models, contact_datas, settings)
 * (constrained-dynamics.hpp:90-147, constrained-dynamics.hxx:193-500) per configuration:
 * the constraint Jacobian and drift of every constraint (Baumgarte-corrected), then
 * [[-mu I, J], [J^T, M]] [-lambda; ddq] = [rhs + mu lambda_prox; tau - nle] solved with
 * max_iter proximal iterations and the reference's stopping test.  Outputs ddq (nv),
 * lambda (data.lambda_c: the constraints' rows in order, at most 12 in all) and, if
 * `iterations` is not NULL, the iteration count of each configuration (settings.iter;
 * host or device memory like the other arrays).  The constraint table and settings are
 * host memory.  FP32: kinematics, CRBA and RNEA in FP32, the constraint rows, the factor
 * and the solve in FP64; FP64: all in double.  Prebuilt topologies only.  A CONTACT_6D
 * placement error is log6 of the relative placement, taken away from rotations by pi
 * (within cbrt(eps) of pi, where the reference switches to its own singular branch, the
 * value is not specified). */
int rbd_constraint_dynamics(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                            int64_t ldq, const void* v, int64_t ldv, const void* tau, int64_t ldtau, int ncontacts,
                            const rbd_rigid_constraint* contacts, const rbd_proximal_settings* settings, void* ddq,
                            int64_t ldddq, void* lambda, int64_t ldlambda, int32_t* iterations, uint32_t flags);

/* computeDelassusMatrix(model, data, q, contact_models, contact_datas, G, mu)
 * (delassus.hpp:41-60, delassus.hxx:118-300) per configuration: the constraints'
 * operational-space inverse inertia J (M + mu J^T J)^-1 J^T, i.e. J M^-1 J^T for mu = 0 --
 * the reference adds a contact inertia of mass mu (rotational inertia mu I for 6-D
 * constraints) at each constraint frame.  J is the constraints' Jacobian as
 * rbd_constraint_dynamics forms it: the reference's Delassus reads joint1's side only
 * (world contacts, joint2 = 0) in the LOCAL frame, where the two agree.  G holds nrows^2
 * rows per configuration (element (i, j) at row i + j nrows), in the input precision;
 * the factor and the products are FP64.  Prebuilt topologies only. */
int rbd_delassus(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                 int ncontacts, const rbd_rigid_constraint* contacts, double mu, void* G, int64_t ldG, uint32_t flags);

/* computeDampedDelassusMatrixInverse(model, data, q, contact_models, contact_datas, Ginv, mu,
 * scaled) (delassus.hpp:84-102, delassus.hxx:800-836) per configuration: (J M^-1 J^T + mu I)^-1
 * (mu > 0), times mu when `scaled` is nonzero.  Layout, precision and constraints as for
 * rbd_delassus. */
int rbd_damped_delassus_inverse(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                                int64_t ldq, int ncontacts, const rbd_rigid_constraint* contacts, double mu, int scaled,
                                void* Ginv, int64_t ldG, uint32_t flags);

/* ContactCholeskyDecomposition::compute(model, data, contact_models, contact_datas, mu) then
 * solve(rhs) (contact-cholesky.hpp:292-400) per configuration: x = K^-1 rhs with the KKT
 * matrix K = [[-mu I, J], [J^T, M]] at q (J the constraints' Jacobian as in
 * rbd_constraint_dynamics, M from CRBA).  rhs and x hold nrows + nv rows per configuration,
 * the constraint rows first.  Solved through the Schur complement J M^-1 J^T + mu I (the
 * reference factors K itself); the factor and the solve are FP64. */
int rbd_contact_cholesky_solve(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q,
                               int64_t ldq, int ncontacts, const rbd_rigid_constraint* contacts, double mu,
                               const void* rhs, int64_t ldrhs, void* x, int64_t ldx, uint32_t flags);

/* Frictional contact problems over Coulomb friction cones (constraints/coulomb-friction-
 * cone.hpp:20-290), batched: `batch` independent problems of `ncones` cones (1..4, problem
 * size n = 3 ncones; each cone's rows are tangent x, tangent y, normal), FP64 only.
 * Per problem: the Delassus matrix G (n^2 rows, element (i, j) at row i + j n), the free
 * velocity g (n rows), the friction coefficients mu (ncones rows).  Layouts and flags as
 * for the other batched arguments (RBD_LAYOUT_BATCH, RBD_HOST_PTRS, RBD_ASYNC).
 * stats: RBD_NCP_STATS rows per problem -- iterations (the reference's getIterationCount),
 * absolute residual, relative residual, converged (1 / 0, the solve's return value), then
 * (ADMM) rho, rho_power, the largest eigenvalue of G and the number of factorizations. */
enum { RBD_NCP_STATS = 8 };

/* PGSContactSolverTpl (pgs-solver.hpp:16-59, contact-solver-base.hpp:28-33); NULL = the
 * reference's defaults (max_it 1000, precisions 1e-6, over_relax 1). */
typedef struct rbd_pgs_settings {
  int32_t max_it;
  double absolute_precision, relative_precision, over_relax;
} rbd_pgs_settings;

/* PGSContactSolverTpl::solve(G, g, cones, x, over_relax) (pgs-solver.hxx:19-141): x (n rows)
 * holds the initial guess and receives the solution. */
int rbd_pgs_solve(rbd_ctx* ctx, int64_t batch, int ncones, const double* G, int64_t ldG, const double* g,
                  int64_t ldg, const double* mu, int64_t ldmu, double* x, int64_t ldx,
                  const rbd_pgs_settings* settings, double* stats, int64_t ldstats, uint32_t flags);

/* ADMMContactSolverTpl (admm-solver.hpp:120-127) parameters with the solver base's
 * (contact-solver-base.hpp:28-33); NULL = the reference's defaults (mu_prox 1e-6, tau 0.5,
 * rho_power 0.2, rho_power_factor 0.05, ratio_primal_dual 10, 20 power iterations,
 * max_it 1000, precisions 1e-6). */
typedef struct rbd_admm_settings {
  double mu_prox, tau, rho_power, rho_power_factor, ratio_primal_dual;
  int32_t max_it_largest_eigen_value, max_it;
  double absolute_precision, relative_precision;
} rbd_admm_settings;

/* ADMMContactSolverTpl::solve(delassus, g, cones, R, primal_guess, dual_guess)
 * (admm-solver.hxx:22-335) with a fresh solver on a fresh DelassusOperatorDense of G per
 * problem (the largest eigenvalue by power iteration from the uniform vector): outputs y
 * (getPrimalSolution), z (getDualSolution) and, if `s` is not NULL, the complementarity
 * shift (getComplementarityShift), n rows each.  R (n rows) is the compliance;
 * primal_guess / dual_guess (n rows, ldguess) may be NULL for none. */
int rbd_admm_solve(rbd_ctx* ctx, int64_t batch, int ncones, const double* G, int64_t ldG, const double* g,
                   int64_t ldg, const double* mu, int64_t ldmu, const double* R, int64_t ldR,
                   const double* primal_guess, const double* dual_guess, int64_t ldguess,
                   const rbd_admm_settings* settings, double* y, double* z, double* s, int64_t ldout, double* stats,
                   int64_t ldstats, uint32_t flags);

/* impulseDynamics (contact-dynamics.hpp:241-300, contact-dynamics.hxx:264-331): the velocity
 * after an impact, dq_after = v_before + M^-1 J^T impulse with
 * impulse = (J M^-1 J^T + inv_damping I)^-1 (-(1 + r_coeff) J v_before), the reference's
 * data.impulse_c.  J as in rbd_constrained_forward_dynamics (1 <= nc <= 12, nc nv rows
 * per configuration); CRBA in the input precision, the factor and the rest in FP64. */
int rbd_impulse_dynamics(rbd_ctx* ctx, const rbd_model* model, int64_t batch, int dtype, const void* q, int64_t ldq,
                         const void* v_before, int64_t ldv, int nc, const void* J, int64_t ldJ, double r_coeff,
                         double inv_damping, void* dq_after, int64_t ldda, void* impulse, int64_t ldimpulse,
                         uint32_t flags);

/* Mixed-robot batch (no reference counterpart: the reference batches one model per
 * call).  Each group is `count` instances of one model, with FP32 device arrays in the
 * batch-fastest layout (element (k, b) at p[k * count + b]); every instance runs RNEA
 * (tau_out from q, v, a) and ABA (ddq_out from q, v, tau).  All groups go to the GPU in
 * one kernel launch over a cost-ordered tile table.  On the first call with a given set
 * of groups the autotuner times the candidate tile widths (64-instance chunks per
 * workgroup, per model) and a per-model launch variant, keeps the fastest and captures
 * it in a HIP graph that later calls with the same arguments replay. */
typedef struct rbd_batch_group {
  const rbd_model* model;
  int64_t count;
  const float* q;
  const float* v;
  const float* a;
  const float* tau;
  float* tau_out;
  float* ddq_out;
} rbd_batch_group;

#define RBD_NO_GRAPH 0x10u    /* rbd_multi_rnea_aba: launch directly, no graph capture/replay */
#define RBD_NO_AUTOTUNE 0x20u /* rbd_multi_rnea_aba: use the default tile widths */

int rbd_multi_rnea_aba(rbd_ctx* ctx, int32_t ngroups, const rbd_batch_group* groups, uint32_t flags);
/* The autotuner's last decision and timings as a JSON string (NUL-terminated, truncated to len). */
int rbd_multi_tuning(rbd_ctx* ctx, char* buf, int32_t len);

/* Learned forward-dynamics surrogate (no reference counterpart).  A multilayer perceptron
 * in -> H -> ... -> H -> nv with SiLU activations, BF16 weights and activations, FP32
 * accumulation (matrix cores), FP32 master weights trained with Adam on an MSE loss in
 * normalised units.  Inputs per configuration: sin and cos of every revolute coordinate,
 * the free-flyer quaternion and its rotation matrix (if the root is a free flyer), v and
 * tau -- 73 features for the 18-DoF quadruped -- each normalised as (x - mean) / std.
 * All batched arrays are FP32 device arrays in the batch-fastest layout (element (k, b) at
 * p[k * ld + b], ld >= batch).  Parameters are exchanged in the packed order: per layer,
 * W (out x in, row major) then b (out). */
typedef struct rbd_mlp rbd_mlp;
int rbd_mlp_create(rbd_ctx* ctx, const rbd_model* model, int32_t hidden, int32_t nhidden, uint64_t seed,
                   rbd_mlp** out);
int rbd_mlp_destroy(rbd_mlp* mlp);
int rbd_mlp_dims(const rbd_mlp* mlp, int32_t* nin, int32_t* nout, int32_t* hidden, int32_t* nhidden,
                 int64_t* nparams);
/* host arrays of nin (x) and nout (y) entries; scales must be positive */
int rbd_mlp_set_normalization(rbd_mlp* mlp, const float* xmean, const float* xstd, const float* ymean,
                              const float* ystd);
/* host array of nparams floats (packed order); set also resets the optimiser state */
int rbd_mlp_get_params(rbd_mlp* mlp, float* params, int64_t nparams);
int rbd_mlp_set_params(rbd_mlp* mlp, const float* params, int64_t nparams);
/* raw (unnormalised) features, device array [batch][nin] (row per configuration) */
int rbd_mlp_features(rbd_ctx* ctx, rbd_mlp* mlp, int64_t batch, const float* q, int64_t ldq, const float* v,
                     int64_t ldv, const float* tau, int64_t ldtau, float* features, uint32_t flags);
/* ddq = surrogate(q, v, tau), no check.  Networks of hidden width 128, 256 or 512 with at
 * most 64 outputs run as one fused kernel (activations kept on chip); RBD_MLP_LAYERED
 * evaluates them layer by layer instead, with the training path's kernels. */
#define RBD_MLP_LAYERED 0x40u
int rbd_mlp_predict(rbd_ctx* ctx, rbd_mlp* mlp, int64_t batch, const float* q, int64_t ldq, const float* v,
                    int64_t ldv, const float* tau, int64_t ldtau, float* ddq, int64_t ldddq, uint32_t flags);
/* one Adam step (lr; beta 0.9, 0.999; eps 1e-8) on the minibatch with targets ddq; the
 * minibatch MSE (normalised units) is added to *loss (a device float) when loss != NULL */
int rbd_mlp_train_step(rbd_ctx* ctx, rbd_mlp* mlp, int64_t batch, const float* q, int64_t ldq, const float* v,
                       int64_t ldv, const float* tau, int64_t ldtau, const float* ddq, int64_t ldddq, float lr,
                       float* loss, uint32_t flags);
/* loss (host float) and its gradient w.r.t. every parameter (host array, packed order),
 * no update */
int rbd_mlp_gradients(rbd_ctx* ctx, rbd_mlp* mlp, int64_t batch, const float* q, int64_t ldq, const float* v,
                      int64_t ldv, const float* tau, int64_t ldtau, const float* ddq, int64_t ldddq, float* grad,
                      int64_t nparams, float* loss);
/* Forward dynamics through the surrogate, checked per configuration by the RNEA residual
 * ||rnea(q, v, ddq) - tau||_2 <= tol ||tau||_2; configurations above it are recomputed with
 * ABA.  *nfallback (host) receives their count; accepted (device, batch bytes: 1 = surrogate
 * kept) and residual (device, batch floats: the surrogate's relative residual) may be
 * NULL.  Synchronous. */
int rbd_mlp_forward_dynamics(rbd_ctx* ctx, rbd_mlp* mlp, int64_t batch, const float* q, int64_t ldq, const float* v,
                             int64_t ldv, const float* tau, int64_t ldtau, float tol, float* ddq, int64_t ldddq,
                             int64_t* nfallback, unsigned char* accepted, float* residual);

/* Seeded FP32 batch inputs generated on the device (batch-fastest layout), draw for draw
 * the host sampler of the package (pinocchio_amd/models.py sample_inputs): per joint in
 * order, free-flyer translation U(-base_range, base_range)^3 and a uniform unit quaternion
 * [x, y, z, w], revolute angles U(-pi, pi); then v ~ N(0, 1), a ~ N(0, 1), tau ~
 * N(0, tau_sigma^2).  a may be NULL. */
int rbd_sample_inputs(rbd_ctx* ctx, const rbd_model* model, int64_t batch, uint64_t seed, double tau_sigma,
                      double base_range, float* q, int64_t ldq, float* v, int64_t ldv, float* a, int64_t lda,
                      float* tau, int64_t ldtau, uint32_t flags);

#ifdef __cplusplus
}
#endif

#endif /* RBD_H_ */
