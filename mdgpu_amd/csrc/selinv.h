// Schedule of the supernodal selected inverse Z = A^{-1} on the pattern of
// L (FP64), computed top-down over the front tree from the stored factor,
// for the reverse mode (gradient of log det A and of solves).
//
// For front f with pivot block J (64 columns at a time, from the last block
// to the first) and the rows R after J (later pivots of f and its update
// rows), with Lh = L(R, J) inv(L(J, J)):
//     Z(R, J) = -Z(R, R) Lh
//     Z(J, J) = inv(L(J,J))^T inv(L(J,J)) - Lh^T Z(R, J)
// (the Takahashi recurrences of Z L = L^{-T}).  Z(R, R) lives in a dense
// symmetric work matrix ZF_f (n_f x n_f): its update-row block Z22_f comes
// from the parent's ZF through the relative map (the reverse of extend-add,
// written into the factorization's update arena at U_f's offset -- the same
// lifetime, level(f) .. level(parent), run backwards), and its pivot blocks
// are filled as they are computed.  The pivot columns of ZF_f are finally
// copied to the Z store, which has the layout of the L store.
#pragma once

#include <string>
#include <vector>

#include "plan.h"

namespace mdgpu {

// buffers addressed by the selected-inverse jobs
enum SelBuf { SB_L = 0, SB_Z, SB_ZF, SB_LH, SB_W, SB_U, SB_N };

// C = alpha op(A) op(B) + beta C on 64x64 tiles; op(X) = X (t = 0) or X^T
// (t = 1); column-major, element offsets into the buffers SelBuf.
struct SelGemm {
    int64_t a_off, b_off, c_off;
    int a_buf, b_buf, c_buf;
    int lda, ldb, ldc;
    int M, N, K;
    int ta, tb;
    double alpha, beta;
    int tiles_m, tile_base, ntiles, pad;
};

enum SelCopyKind {
    SC_INIT = 0,   // ZF_f(c+i, c+j) = Z22_f(i, j)              columns [j0, j1) of Z22
    SC_TRANS,      // ZF_f(J, i) = ZF_f(i, J), i in rows [j0, j1) of R, block J = [b0, b0+b)
    SC_OUT,        // Z store(:, j) = ZF_f(:, j)                 pivot columns [j0, j1)
    SC_PUSH        // Z22_f(i, j) = ZF_parent(rel(i), rel(j))    columns [j0, j1) of Z22_f
};
struct SelCopy {
    int front;
    int j0, j1;
    int b0, b;
    int pad;
};

enum SelStepKind { SS_GEMM = 0, SS_COPY };
struct SelStep {
    int kind;
    int copy_kind;
    int first, count;   // slice of gemm or copy
    int nblocks;        // workgroups to launch
    int level;
};

struct SelPlan {
    std::vector<int64_t> zf_ofs, lh_ofs;   // per front, within the per-level regions
    int64_t zf_size = 1, lh_size = 1;      // doubles (peak over levels)
    std::vector<SelGemm> gemm;
    std::vector<SelCopy> copy;
    std::vector<SelStep> steps;
    double flops = 0;
};

std::string build_selinv_plan(const Plan& P, SelPlan& Q);
std::string validate_selinv_plan(const Plan& P, const SelPlan& Q);

}  // namespace mdgpu
