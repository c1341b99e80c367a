// Krylov control block and the scalar recurrences of CG (Cg.cuh:123-189)
// and BiCGStab (BiCgStab.cuh:144-247), evaluated on the device by the
// finalising kernel of each reduction.  Shared by the single and the
// multi-right-hand-side solvers (one control block per right-hand side).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

namespace tf {

struct KCtl {
    double normb, tol, res, rho, rho_prev, alpha, beta, omega, den, ns;
    int status;  // -1 running, otherwise the status handed to finish()
    int iter;
    int fresh;
    int check;  // true-residual refresh requested
    int early;  // BiCGStab: ||s|| <= tol in this iteration
    int refresh;
};

enum { kRunning = -1, kConverged = 0, kMaxIt = 1, kBreakdown = 2, kNonFinite = 3 };

// Finalisers run by thread 0 of the last block of a reduction.
enum Fin { kFinRhoCg, kFinCurvCg, kFinUpdCg, kFinTrue, kFinRhoBi, kFinDenBi, kFinSBi, kFinUpdBi };

__device__ inline void finalize(int mode, double total, KCtl* c) {
    switch (mode) {
    case kFinRhoCg: {  // Cg.cuh:139-160
        c->iter += 1;
        const double rho = total;
        c->rho = rho;
        if (!isfinite(rho)) c->status = kNonFinite;
        else if (rho <= 0) c->status = kBreakdown;
        else if (!c->fresh) {
            c->beta = rho / c->rho_prev;
            if (!isfinite(c->beta)) c->status = kNonFinite;
        }
        break;
    }
    case kFinCurvCg: {  // Cg.cuh:161-172
        const double curv = total;
        if (!isfinite(curv)) c->status = kNonFinite;
        else if (curv <= 0) c->status = kBreakdown;
        else {
            c->alpha = c->rho / curv;
            if (!isfinite(c->alpha)) c->status = kNonFinite;
        }
        c->fresh = 0;
        break;
    }
    case kFinUpdCg: {  // Cg.cuh:173-187
        const double res = sqrt(total);
        c->res = res;
        if (!isfinite(res)) c->status = kNonFinite;
        else {
            c->check = (res <= c->tol || c->iter % c->refresh == 0) ? 1 : 0;
            c->rho_prev = c->rho;
        }
        break;
    }
    case kFinTrue: {  // true-residual refresh, Cg.cuh:178-185 / BiCgStab.cuh:202-211,235-243
        const double res = sqrt(total);
        c->res = res;
        c->check = 0;
        if (!isfinite(res) || res <= c->tol) c->status = kMaxIt;
        else c->fresh = 1;
        break;
    }
    case kFinRhoBi: {  // BiCgStab.cuh:160-180
        c->iter += 1;
        c->early = 0;
        const double rho = total;
        c->rho = rho;
        if (!isfinite(rho)) c->status = kNonFinite;
        else if (rho == 0) c->status = kBreakdown;
        else if (!c->fresh) {
            c->beta = (rho / c->rho_prev) * (c->alpha / c->omega);
            if (!isfinite(c->beta)) c->status = kNonFinite;
        }
        break;
    }
    case kFinDenBi: {  // BiCgStab.cuh:182-195
        const double den = total;
        c->den = den;
        if (!isfinite(den)) c->status = kNonFinite;
        else if (den == 0) c->status = kBreakdown;
        else {
            c->alpha = c->rho / den;
            if (!isfinite(c->alpha)) c->status = kNonFinite;
        }
        c->fresh = 0;
        break;
    }
    case kFinSBi: {  // BiCgStab.cuh:196-212
        const double ns = sqrt(total);
        c->ns = ns;
        if (!isfinite(ns)) c->status = kNonFinite;
        else c->early = ns <= c->tol ? 1 : 0;
        break;
    }
    case kFinUpdBi: {  // BiCgStab.cuh:229-245
        const double res = sqrt(total);
        c->res = res;
        if (!isfinite(res)) c->status = kNonFinite;
        else {
            c->check = (res <= c->tol || c->iter % c->refresh == 0) ? 1 : 0;
            c->rho_prev = c->rho;
        }
        break;
    }
    default:
        break;
    }
}

__device__ __forceinline__ bool stopped(const KCtl* c) { return c->status != kRunning; }

// Iterations to launch before the host next reads the control blocks of
// `count` systems that have done `done` iterations from initial residuals
// res0: one less than the most iterations any running system's mean residual
// reduction so far says it still needs, at least 1, at most 8.  An iteration
// launched after a system has stopped leaves it unchanged (its Krylov
// kernels return on the stopped status), so the batch never changes results.
inline int krylov_batch(const KCtl* h, const double* res0, int count, int done) {
    if (done < 2) return 1;
    double most = 0;
    for (int s = 0; s < count; ++s) {
        if (h[s].status != kRunning) continue;
        if (!(h[s].res > 0 && h[s].res < res0[s])) return 1;
        const double rate = std::pow(h[s].res / res0[s], 1.0 / done);
        const double remain = std::log(h[s].tol / h[s].res) / std::log(rate);
        if (!std::isfinite(remain)) return 1;
        most = std::max(most, remain);
    }
    return most > 2 ? std::min(8, int(most) - 1) : 1;
}


}  // namespace tf
