// yz_mlp.h -- the learned preconditioner of the film-height step.
//
// A per-cell MLP, 8 -> 64 -> 64 -> 13 with ReLU, maps the features of a
// cell (the stiffness and shape of its row-scaled IA row, the film height
// and the concentration) to the 13 weights of its row of an approximate
// inverse M of IA: (M v)_i = sum_k m_ik v_{col(i,k)}, the same stencil as
// IA. M is linear in the vector it is applied to, so BiCGSTAB stays valid.
// The output is scaled by 1 / (1 + sum_k!=c |IA_ik|) so that the network
// learns O(1) numbers whatever the stiffness.
//
// Training minimises the probe loss L = (1/PN) sum_p ||v_p - IA M v_p||^2
// (a stochastic estimate of ||I - IA M||_F^2 / N) on the matrices the
// solver itself assembles. Its gradient with respect to m_ik is
// -2/(PN) sum_p (IA^T e_p)_i v_p,col(i,k) with e_p = v_p - IA M v_p; the
// network's gradient follows by back-propagation. Forward and backward run
// on the MFMA units in BF16 with FP32 accumulation
// (__builtin_amdgcn_mfma_f32_16x16x16bf16_1k, one wave per 16 cells); the
// FP32 master weights are updated with Adam.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "yz_solver.h"

namespace yz {

// packed FP32 parameters
enum {
    MLP_IN = 8, MLP_H = 64, MLP_OUT = 13,
    MLP_W1 = 0,                      // W1[64][16] (columns 8..15 zero)
    MLP_B1 = MLP_W1 + 64 * 16,       // b1[64]
    MLP_W2 = MLP_B1 + 64,            // W2[64][64]
    MLP_B2 = MLP_W2 + 64 * 64,       // b2[64]
    MLP_W3 = MLP_B2 + 64,            // W3[16][64] (rows 13..15 zero)
    MLP_B3 = MLP_W3 + 16 * 64,       // b3[16]
    MLP_NP = MLP_B3 + 16
};

struct Mlp {
    int N = 0;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    float *partial = nullptr;        // per-wave partial gradients
    int nwaves = 0;
    float *G = nullptr;              // 13 x N: dL/dm
    float *M = nullptr;              // 13 x N: scratch stencil during training
    float *v = nullptr, *u = nullptr, *e = nullptr, *w = nullptr;
    double *loss_part = nullptr, *loss_dev = nullptr, *loss_host = nullptr;
    int adam_t = 0;
    int trained_passes = 0;
    double loss_first = 0, loss_last = 0, train_ms = 0;
    long long train_iters = 0;
};

int mlp_init(Mlp &m, int N, hipStream_t st);
void mlp_free(Mlp &m);
// M (13 x N planes) from the features and the row scale
int mlp_infer(Mlp &m, int N, const float *feat, const float *mscale, float *M, hipStream_t st);
// the raw outputs Y (13 x N planes, FP16): M = diag(mscale) Y
int mlp_infer_half(Mlp &m, int N, const float *feat, __half *Y, hipStream_t st);
// loss and gradient of the probe loss on operator A (S.op0, S.A) for the
// current parameters: fills m.grads, returns the loss
int mlp_loss_grad(Mlp &m, const SolverT<float> &S, const float *feat, const float *mscale, int nprobe,
                  unsigned seed, double *loss, hipStream_t st);
// iters Adam steps on the probe loss
int mlp_train(Mlp &m, const SolverT<float> &S, const float *feat, const float *mscale, int iters, double lr,
              int nprobe, unsigned seed, double *loss_first, double *loss_last, hipStream_t st);

}  // namespace yz
