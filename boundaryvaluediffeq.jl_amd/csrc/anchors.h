// anchors.h -- the learned initial guess between anchor points.
//
// The warm-start network is evaluated at K equally spaced anchors of [t0, t1]
// per instance; each anchor record holds (y1, y2, f1, f2): the network's state
// and the problem's own slope f(t, y) there.  The guess at a mesh node is the
// cubic Hermite interpolant of the two surrounding anchors (M = 2).
#pragma once

#include <hip/hip_runtime.h>

#include "mirk_wave.h"

__device__ inline double anchor_t(const BatchDev &B, int b, int j, int K) {
    return j == K - 1 ? B.t1[b] : B.t0[b] + (B.t1[b] - B.t0[b]) * ((double)j / (double)(K - 1));
}

// index j of the anchor interval [j, j+1] that holds t
__device__ inline int anchor_interval(int K, double t0, double t1, double t) {
    const double sc = (t - t0) / (t1 - t0) * (double)(K - 1);
    int j = (int)floor(sc);
    return j < 0 ? 0 : (j > K - 2 ? K - 2 : j);
}

// y[0..1] at t from the two anchor records rec[0..3], rec[4..7] of its interval
__device__ inline void anchor_eval(const double *rec, int K, double t0, double t1, double t, double *y) {
    const double sc = (t - t0) / (t1 - t0) * (double)(K - 1);
    const int j = anchor_interval(K, t0, t1, t);
    const double th = sc - (double)j, h = (t1 - t0) / (double)(K - 1);
    const double *a = rec, *c = rec + 4;
    const double th2 = th * th, th3 = th2 * th;
    const double h00 = 2 * th3 - 3 * th2 + 1, h10 = th3 - 2 * th2 + th, h01 = -2 * th3 + 3 * th2, h11 = th3 - th2;
    y[0] = h00 * a[0] + h10 * h * a[2] + h01 * c[0] + h11 * h * c[2];
    y[1] = h00 * a[1] + h10 * h * a[3] + h01 * c[1] + h11 * h * c[3];
}

// y[0..1] at node time t of instance b from its anchor records anc[b][K][4]
__device__ inline void anchor_hermite(const double *anc, int K, int b, double t0, double t1, double t, double *y) {
    anchor_eval(anc + ((size_t)b * K + anchor_interval(K, t0, t1, t)) * 4, K, t0, t1, t, y);
}
