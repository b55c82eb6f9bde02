// singularity_amd/csrc/adjoint.hpp — host-visible declarations of the
// differentiable render (forward with checkpoints + discrete adjoint).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "diff_scene.hpp"

namespace sing {

// Scene in every precision the kernels need; lives in device memory.
struct DiffSet {
    DiffScene<float> sf;
    DiffScene<double> sd;
    DiffScene<Dual<2, float>> tf;   // tangents wrt (a/M, inclination)
    DiffScene<Dual<2, double>> td;
    double tscale;                  // peak-temperature scale parameter
    double T0;                      // unscaled peak temperature (K)
};
// (device memory holds one DiffSet per evaluation of a call)

// Per-pixel device buffers of the differentiable frames of one call.  A call
// evaluates n_eval parameter points of one image size; their pixels are laid
// out one image after another (q = evaluation * Pimg + pixel), and every
// per-pixel array below is indexed by q with P = n_eval * Pimg entries.
// Blocks of the first stage of the loss/gradient reduction (per evaluation).
constexpr std::uint32_t kReduceBlocks = 256;
// reduced per evaluation: loss, the 5 gradient entries, the RK4 step total
constexpr std::uint32_t kReduceVals = 7;
// the reduction's output: those sums, then the FP64 list's length
constexpr std::uint32_t kDiffOut = kReduceVals + 1;

struct DiffBuffers {
    double* cp;             // [ncp][5][P] states every C steps (double storage)
    double* cross;          // [8][7][P] step, alpha, emission rgb, r_cross, phi
    std::uint32_t* meta;    // [P][4] steps, crossings, termination, precision (1 = FP64)
    double* sky;            // [P][3] escape colour behind the crossings
    double* EL;             // [2][P] E and L_z of the forward (the backward recomputes with these)
    double* gI;             // [P][3] dLoss/dRadiance
    float* hdr;             // [P][3] radiance
    double* loss;           // [P]
    double* grad;           // [P][5]
    double* scratch;        // [C+1][5][P] recomputed segment states
    const float* target;    // [Pimg][3] shared by the evaluations, or [n_eval][Pimg][3] (tstride = Pimg)
    const float* weight;    // [Pimg], [n_eval][Pimg] or null
    std::uint32_t tstride;  // pixels between evaluations' targets and weights (0: one shared image)
    // [Pimg][3] or null: a caller's dLoss/dRadiance (sing_diff_vjp); the loss
    // is then the pairing sum_p cot_p . I_p, whose gradient is the
    // vector-Jacobian product, and target/weight are not read
    const float* cot;
    std::uint32_t* list;    // [n_eval][Pimg] pixels of the FP64 pass (classified hard before the march)
    std::uint32_t* count;   // [n_eval] lengths of list
    std::uint32_t* list2;   // [n_eval][Pimg] pixels handed to FP64 after the FP32 march (grazing, near-hole emission)
    std::uint32_t* count2;  // [n_eval] lengths of list2
    std::uint32_t* steps;   // [P] RK4 steps of the forward
    double* partial;        // [n_eval][kReduceBlocks][kReduceVals] per-block sums of loss, gradient, steps
    std::uint32_t P, ncp, C;
    std::uint32_t Pimg;     // pixels of one image (P = n_eval * Pimg)
    std::uint32_t n_eval;
};

// out: n_eval x kDiffOut doubles (per evaluation: loss, gradient, step total,
// then the length of its FP64 list).
hipError_t launch_diff(const DiffSet* ds_dev, const DiffBuffers& b, std::uint32_t precision, bool backward,
                       double* out, hipStream_t stream, hipStream_t stream2, hipEvent_t fwd_end,
                       const hipEvent_t* join);

}  // namespace sing
