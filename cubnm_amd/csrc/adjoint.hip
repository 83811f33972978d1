// Gradients of a BOLD-based loss with respect to rWW parameters through the
// simulation (no counterpart in the reference, which fits with CMA-ES and
// grid search only): a hand-written discrete adjoint of the Euler-Maruyama
// rWW steps, the coupling, the Balloon-Windkessel steps and the BOLD
// read-out, with the noise held fixed.
//
// Forward (`rww_fwd_kernel`): one workgroup per simulation, one lane per
// node, the same arithmetic as the feature-complete integrator (coupling
// sum_k SC[k][j] S_E(k) in node order, RWW::step, bw_step, bold_signal), so
// its BOLD equals run_simulations' bit for bit.  Every `ck` BW steps it
// stores (S_E, S_I, x, f, nu, q) of every node: a checkpoint.
//
// Backward (`rww_adj_kernel`): segments are processed last to first.  Each
// segment is recomputed from its checkpoint, keeping S_E, S_I and the
// coupling input g of every inner step and the BW state of every BW step in
// a per-simulation scratch; then the adjoint runs backwards through it:
//   BOLD:  y = V0 (k1 (1-q) + k2 (1-q/nu) + k3 (1-nu))
//   BW:    x' = x + dt (u - kappa x - gamma (f-1)), f' = f + dt x',
//          nu' = nu + dt/tau (f - nu^a), q' = q + dt/tau (f (1-(1-rho)^(1/f))/rho - nu'^a q / nu')
//   rWW:   I_E = w_E I_0 + w_p J S_E + G J g - w_IE S_I, I_I = w_I I_0 + J S_E - w_II S_I,
//          S_E* = S_E + noise + dt gE (1-S_E) H_E(I_E) - dt/tE S_E, S_E' = clip(S_E*, 0, 1),
//          S_I* = S_I + noise + dt gI H_I(I_I) - dt/tI S_I,         S_I' = clip(S_I*, 0, 1),
//          H(x) = u / (1 - exp(-d u)), u = a x - b;
//   coupling: the adjoint of g_j = sum_k SC[k][j] S_E(k) is
//          lambda_SE(k) += sum_j SC[k][j] e_j with e_j = lambda(I_E(j)) G J_j,
//          read from an LDS copy of SC (rows padded by one double).
// Parameter gradients accumulate dI_E/dG = J g, dI_E/dw_p = J S_E,
// dI_E/dJ = w_p S_E + G g, dI_I/dJ = S_E.  Clipping passes no gradient where
// it is active.  Memory: checkpoints 6 N doubles per ck BW steps; scratch
// (3 N inner_it + 5 N) doubles per BW step of one segment, per simulation.
#include "engine.hpp"
#include "../../include/cubnm_amd.h"
#include <string>

namespace cubnm {

BWConsts current_bw();
void set_last_error(const std::string& e);

namespace {

struct AdjArgs {
    int n_sims, N, bw_it, inner_it, TR_iters, bold_len, ck, n_ck, noise_bw_it;
    const double *SC, *G, *wp, *JN, *wIE, *sigma;   // SC [N][N] shared; G [S]; regional [S][N]
    const double* noise;                            // [noise_bw_it*inner_it][N][2]
    const int *sh_nodes, *sh_ts;                    // [repeats][N], [repeats][noise_bw_it]
    double* bold;                                   // [S][bold_len][N] (forward output)
    double* ckpt;                                   // [S][n_ck][6][N]
    const double* dbold;                            // [S][bold_len][N] (backward input)
    double* scratch;                                // [S][ck][inner_it*3 + 5][N]
    double *gG, *gwp, *gJN;                         // [S], [S][N], [S][N]
    int sc_lds;                                     // SC staged in LDS (padded rows)
    int np, ks;                                     // lanes per k partition, partitions
};

// Shuffled segment step and node of BW step bw_i (once per BW step), then the
// two noise values of one inner step (the reference's index, bnm.cu:300-312).
__device__ inline void noise_shuffle(const AdjArgs& a, int bw_i, int j, int* sh_t, int* sh_j) {
    const int rep = bw_i / a.noise_bw_it;
    *sh_t = a.sh_ts[(bw_i % a.noise_bw_it) + rep * a.noise_bw_it];
    *sh_j = a.sh_nodes[rep * a.N + j];
}
__device__ inline void noise_load(const AdjArgs& a, int sh_t, int sh_j, int inner_i, double* n0, double* n1) {
    const long long idx = ((long long)(sh_t * a.inner_it + inner_i) * a.N + sh_j) * 2;
    *n0 = a.noise[idx];
    *n1 = a.noise[idx + 1];
}

// H(x) = u / (1 - exp(-d u)) with u = a x - b, and its derivative
// dH/dx = a (den - u d e) / den^2 with e = exp(-d u), den = 1 - e: one
// exponential for both, divisions by the reciprocal sequence (fast_div).
__device__ inline double rate_and_slope(double a, double b, double d, double x, double* slope) {
    const double u = a * x - b;
    const double e = exp_cw(-d * u);
    const double den = 1.0 - e;
    *slope = a * fast_div(den - u * d * e, den * den);
    return fast_div(u, den);
}

// Adjoint state of one node (lambda of S_E, S_I and of the BW variables) and
// its parameter-gradient accumulators.
struct RevState {
    double lSE = 0, lSI = 0, lx = 0, lf = 0, ln = 0, lq = 0;
    double gG = 0, gwp = 0, gJN = 0;
};

// Adjoint of BW step b (pre-step state x, f, nu, q, input u) and of the BOLD
// read-out that follows it when b ends a TR; on entry the lambdas refer to
// the post-step state, on exit to the pre-step state, and lSE has received
// the input's share.
__device__ inline void bw_adjoint(const AdjArgs& a, const BWConsts& bw, int b, double x, double f, double nu,
                                  double q, double u, const double* dbold, int jj, double lnr1, RevState& r) {
    const double a_dt = bw.dt, a_it = bw.dt_itau, ra = bw.ialpha, r1 = bw.oneminrho;
    double x1 = x, f1 = f, nu1 = nu, q1 = q;
    bw_step(bw, u, x1, f1, nu1, q1);
    if ((b + 1) % a.TR_iters == 0) {
        const int vol = (b + 1) / a.TR_iters - 1;
        if (vol < a.bold_len) {
            const double dy = dbold[(long long)vol * a.N + jj];
            r.lq += dy * bw.V_0 * (-bw.k1 - bw.k2 / nu1);
            r.ln += dy * bw.V_0 * (bw.k2 * q1 / (nu1 * nu1) - bw.k3);
        }
    }
    const double nu1a = pow(nu1, ra), nua = pow(nu, ra), r1f = pow(r1, 1.0 / f);
    const double lq_pre = r.lq * (1.0 - a_it * nu1a / nu1);
    const double ln1 = r.ln + r.lq * (-a_it * q * (ra - 1.0) * nu1a / (nu1 * nu1));
    double lf_pre = r.lq * a_it / bw.rho * (1.0 - r1f + r1f * lnr1 / f);
    const double ln_pre = ln1 * (1.0 - a_it * ra * nua / nu);
    lf_pre += ln1 * a_it;
    const double lx1 = r.lx + r.lf * a_dt;
    lf_pre += r.lf - lx1 * a_dt * bw.y;
    const double lx_pre = lx1 * (1.0 - a_dt * bw.kappa);
    const double lu = lx1 * a_dt;
    r.lx = lx_pre; r.lf = lf_pre; r.ln = ln_pre; r.lq = lq_pre;
    r.lSE += lu;
}

// First half of an inner rWW step's adjoint, from the step's S_E, S_I,
// coupling input g and noise (v = {S_E, S_I, g, n0, n1}): the clipping masks
// applied to lambda (muE, muI), the sensitivities of the two populations'
// inputs (eE, eI), the rate r_E, and the parameter gradients.
__device__ inline void rww_adjoint_pre(const RWW::C& c, const double* v, double G, double wp, double JN, double wIE,
                                       double sig, RevState& r, double& eE, double& eI, double& muE, double& muI,
                                       double& rE) {
    const double SE = v[0], SI = v[1], g = v[2], n0 = v[3], n1 = v[4];
    const double IE = c.w_E__I_0 + wp * JN * SE + g * G * JN - wIE * SI;
    const double II = c.w_I__I_0 + JN * SE - c.w_II * SI;
    double hE, hI;
    rE = rate_and_slope(c.a_E, c.b_E, c.d_E, IE, &hE);
    const double rI = rate_and_slope(c.a_I, c.b_I, c.d_I, II, &hI);
    const double SEs = SE + n0 * sig * c.sqrt_dt + c.dt_gamma_E * ((1 - SE) * rE) - c.dt_itau_E * SE;
    const double SIs = SI + n1 * sig * c.sqrt_dt + c.dt_gamma_I * rI - c.dt_itau_I * SI;
    muE = (SEs > 0.0 && SEs < 1.0) ? r.lSE : 0.0;
    muI = (SIs > 0.0 && SIs < 1.0) ? r.lSI : 0.0;
    eE = muE * c.dt_gamma_E * (1 - SE) * hE;
    eI = muI * c.dt_gamma_I * hI;
    r.gG += eE * JN * g;
    r.gwp += eE * JN * SE;
    r.gJN += eE * (wp * SE + G * g) + eI * SE;
}

// Second half: lambda of the pre-step S_E and S_I, with coup = sum_j SC[k][j] e_j.
__device__ inline void rww_adjoint_post(const RWW::C& c, double eE, double eI, double muE, double muI, double rE,
                                        double wp, double JN, double wIE, double coup, RevState& r) {
    r.lSE = muE * (1.0 - c.dt_gamma_E * rE - c.dt_itau_E) + eE * wp * JN + eI * JN + coup;
    r.lSI = muI * (1.0 - c.dt_itau_I) - eE * wIE - eI * c.w_II;
}

// Writes the reverse lanes' gradients; the G gradient is summed over the
// first `lanes` threads of the block (a multiple of 64).
__device__ inline void store_gradients(const AdjArgs& a, int sim, int j, bool act, int lanes, const RevState& r,
                                       double* red) {
    if (act) {
        a.gwp[(long long)sim * a.N + j] = r.gwp;
        a.gJN[(long long)sim * a.N + j] = r.gJN;
    }
    double v = act ? r.gG : 0.0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && (int)threadIdx.x < lanes) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < lanes / 64; w++) t += red[w];
        a.gG[sim] = t;
    }
}
}  // namespace

// MAXT: launch bound (256 for up to 256 nodes, which lifts the register
// budget from 128 to 256 VGPRs: the backward pass keeps one step of
// prefetched scratch, noise and BW state in registers)
template <int MAXT>
__global__ void __launch_bounds__(MAXT) rww_fwd_kernel(AdjArgs a, RWW::C c, BWConsts bw) {
    extern __shared__ double lds[];
    const int sim = blockIdx.x, N = a.N, np = a.np, ks = a.ks;
    // ks partitions of np lanes; partition p sums sources [k0, k1), partition 0
    // owns the node state and adds the others' partial sums in partition order
    const int part = threadIdx.x / np, j = threadIdx.x - part * np;
    const bool own = part == 0;
    const bool act = own && j < N;
    const int jj = j < N ? j : 0;
    const int kc = (N + ks - 1) / ks, k0 = min(N, part * kc), k1 = min(N, k0 + kc);
    double* conn = lds;                   // [N]
    double* redk = lds + N;               // [ks-1][np] partial sums
    double* scs = redk + (ks - 1) * np;   // [N][N+1] SC copy when a.sc_lds
    const int ldsc = N + 1;
    if (a.sc_lds)
        for (int i = threadIdx.x; i < N * N; i += blockDim.x) scs[(i / N) * ldsc + i % N] = a.SC[i];
    const double gp[1] = {a.G[sim]};
    const double rp[4] = {a.wp[(long long)sim * N + jj], a.JN[(long long)sim * N + jj],
                          a.wIE[(long long)sim * N + jj], a.sigma[(long long)sim * N + jj]};
    double s[RWW::n_state] = {0, 0, 0, 0, 0, 0};
    RWW::init(c, s, rp);
    double bx = 0.0, bf = 1.0, bn = 1.0, bq = 1.0;
    if (act) conn[j] = s[4];
    double* bold = a.bold + (long long)sim * a.bold_len * N;
    int bold_i = 0;
    int sh_t = 0, sh_j = 0;
    double pn0 = 0.0, pn1 = 0.0;  // noise of the next inner step
    if (own) {
        noise_shuffle(a, 0, jj, &sh_t, &sh_j);
        noise_load(a, sh_t, sh_j, 0, &pn0, &pn1);
    }
    for (int bw_i = 0; bw_i < a.bw_it; bw_i++) {
        if (bw_i % a.ck == 0 && act) {
            double* ck = a.ckpt + ((long long)sim * a.n_ck + bw_i / a.ck) * 6 * N;
            ck[0 * N + j] = s[4]; ck[1 * N + j] = s[5];
            ck[2 * N + j] = bx; ck[3 * N + j] = bf; ck[4 * N + j] = bn; ck[5 * N + j] = bq;
        }
        // shuffles of the next BW step, for the noise prefetch of its first step
        int nsh_t = sh_t, nsh_j = sh_j;
        if (own && bw_i + 1 < a.bw_it) noise_shuffle(a, bw_i + 1, jj, &nsh_t, &nsh_j);
        for (int inner_i = 0; inner_i < a.inner_it; inner_i++) {
            // this step's noise was loaded one step ahead; load the next
            // step's now, so that its latency overlaps a whole step (the
            // barriers order LDS only)
            const double n0 = pn0, n1 = pn1;
            if (own) {
                if (inner_i + 1 < a.inner_it) noise_load(a, sh_t, sh_j, inner_i + 1, &pn0, &pn1);
                else if (bw_i + 1 < a.bw_it) noise_load(a, nsh_t, nsh_j, 0, &pn0, &pn1);
            }
            lds_barrier();
            double g = 0.0;
            if (a.sc_lds) {
#pragma unroll 8
                for (int k = k0; k < k1; k++) g += scs[k * ldsc + jj] * conn[k];
            } else {
#pragma unroll 8
                for (int k = k0; k < k1; k++) g += a.SC[(long long)k * N + jj] * conn[k];
            }
            if (ks > 1) {
                if (!own) redk[(part - 1) * np + j] = g;
                lds_barrier();
                if (own)
                    for (int p = 1; p < ks; p++) g += redk[(p - 1) * np + j];
            }
            if (own) RWW::step(c, s, gp, rp, g, n0, n1);
            lds_barrier();
            if (act) conn[j] = s[4];
        }
        sh_t = nsh_t; sh_j = nsh_j;
        if (own) bw_step(bw, s[4], bx, bf, bn, bq);
        if ((bw_i + 1) % a.TR_iters == 0) {
            if (act && bold_i < a.bold_len) bold[(long long)bold_i * N + j] = bold_signal(bw, bn, bq);
            bold_i++;
        }
    }
}

template <int MAXT>
__global__ void __launch_bounds__(MAXT) rww_adj_kernel(AdjArgs a, RWW::C c, BWConsts bw) {
    extern __shared__ double lds[];
    const int sim = blockIdx.x, N = a.N, np = a.np, ks = a.ks;
    // k partitions as in the forward kernel: partition 0 owns the node state
    // and the adjoint state; every partition sums its range of sources in the
    // coupling (recompute) and in the coupling's transpose (backward)
    const int part = threadIdx.x / np, j = threadIdx.x - part * np;
    const bool own = part == 0;
    const bool act = own && j < N;
    const int jj = j < N ? j : 0;
    const int kc = (N + ks - 1) / ks, k0 = min(N, part * kc), k1 = min(N, k0 + kc);
    const int ldsc = N + 1;
    double* conn = lds;                 // [N]
    double* ev = lds + N;               // [N]   e_j of the coupling adjoint
    double* red = lds + 2 * N;          // [16]  block reduction
    double* redk = red + 16;            // [ks-1][np] partial sums
    double* scs = redk + (ks - 1) * np; // [N][N+1] SC copy when a.sc_lds
    if (a.sc_lds)
        for (int i = threadIdx.x; i < N * N; i += blockDim.x) scs[(i / N) * ldsc + i % N] = a.SC[i];
    const double G = a.G[sim];
    const double wp = a.wp[(long long)sim * N + jj], JN = a.JN[(long long)sim * N + jj];
    const double wIE = a.wIE[(long long)sim * N + jj], sig = a.sigma[(long long)sim * N + jj];
    const double gp[1] = {G};
    const double rp[4] = {wp, JN, wIE, sig};
    const int per_bw = a.inner_it * 3 + 5;
    double* scr = a.scratch + (long long)sim * a.ck * per_bw * N;
    const double* dbold = a.dbold + (long long)sim * a.bold_len * N;
    const double lnr1 = log(bw.oneminrho);
    RevState r;  // adjoint state and gradient accumulators
    for (int seg = a.n_ck - 1; seg >= 0; seg--) {
        const int b0 = seg * a.ck, b1 = min(a.bw_it, b0 + a.ck);
        // ---- recompute the segment from its checkpoint
        const double* ck = a.ckpt + ((long long)sim * a.n_ck + seg) * 6 * N;
        double s[RWW::n_state] = {0, 0, 0, 0, 0, 0};
        double bx = 0, bf = 1, bn = 1, bq = 1;
        if (own) {
            s[4] = ck[0 * N + jj]; s[5] = ck[1 * N + jj];
            bx = ck[2 * N + jj]; bf = ck[3 * N + jj]; bn = ck[4 * N + jj]; bq = ck[5 * N + jj];
        }
        __syncthreads();
        if (act) conn[j] = s[4];
        // noise loaded one inner step ahead, as in the forward kernel
        int sh_t = 0, sh_j = 0;
        double pn0 = 0.0, pn1 = 0.0;
        if (own) {
            noise_shuffle(a, b0, jj, &sh_t, &sh_j);
            noise_load(a, sh_t, sh_j, 0, &pn0, &pn1);
        }
        for (int b = b0; b < b1; b++) {
            double* sb = scr + (long long)(b - b0) * per_bw * N;
            int nsh_t = sh_t, nsh_j = sh_j;
            if (own && b + 1 < b1) noise_shuffle(a, b + 1, jj, &nsh_t, &nsh_j);
            for (int i = 0; i < a.inner_it; i++) {
                const double n0 = pn0, n1 = pn1;
                if (own) {
                    if (i + 1 < a.inner_it) noise_load(a, sh_t, sh_j, i + 1, &pn0, &pn1);
                    else if (b + 1 < b1) noise_load(a, nsh_t, nsh_j, 0, &pn0, &pn1);
                }
                lds_barrier();
                double g = 0.0;
                if (a.sc_lds) {
#pragma unroll 8
                    for (int k = k0; k < k1; k++) g += scs[k * ldsc + jj] * conn[k];
                } else {
#pragma unroll 8
                    for (int k = k0; k < k1; k++) g += a.SC[(long long)k * N + jj] * conn[k];
                }
                if (ks > 1) {
                    if (!own) redk[(part - 1) * np + j] = g;
                    lds_barrier();
                    if (own)
                        for (int p = 1; p < ks; p++) g += redk[(p - 1) * np + j];
                }
                // each lane reads back only its own scratch entries
                if (act) { sb[(3 * i) * N + j] = s[4]; sb[(3 * i + 1) * N + j] = s[5]; sb[(3 * i + 2) * N + j] = g; }
                if (own) RWW::step(c, s, gp, rp, g, n0, n1);
                lds_barrier();
                if (act) conn[j] = s[4];
            }
            double* bwst = sb + (long long)3 * a.inner_it * N;
            if (act) { bwst[0 * N + j] = bx; bwst[1 * N + j] = bf; bwst[2 * N + j] = bn; bwst[3 * N + j] = bq; bwst[4 * N + j] = s[4]; }
            if (own) bw_step(bw, s[4], bx, bf, bn, bq);
            sh_t = nsh_t; sh_j = nsh_j;
        }
        // ---- backwards through the segment.  Step t = (b - b0) * inner_it + i
        // runs from T-1 down to 0; the scratch entries and noise of step t-2
        // are loaded while step t is processed (two steps ahead: these are
        // HBM reads of data written a segment earlier), the BW-step state one
        // BW step ahead; the barriers order LDS only.
        const int T = (b1 - b0) * a.inner_it;
        int fb = -1, fsh_t = 0, fsh_j = 0;  // block and noise shuffles of the last fetch
        auto fetch = [&](int t, double* v) {  // v = S_E, S_I, g, noise 0, noise 1 of step t
            const int bb = b0 + t / a.inner_it, ii = t % a.inner_it;
            if (bb != fb) { noise_shuffle(a, bb, jj, &fsh_t, &fsh_j); fb = bb; }
            const double* sbx = scr + (long long)(bb - b0) * per_bw * N;
            v[0] = sbx[(3 * ii) * N + jj]; v[1] = sbx[(3 * ii + 1) * N + jj]; v[2] = sbx[(3 * ii + 2) * N + jj];
            noise_load(a, fsh_t, fsh_j, ii, &v[3], &v[4]);
        };
        double p1[5] = {0, 0, 0, 0, 0}, p2[5] = {0, 0, 0, 0, 0};  // steps t and t-1
        double pbx = 0, pbf = 1, pbn = 1, pbq = 1, pbu = 0;
        if (own) {
            fetch(T - 1, p1);
            if (T >= 2) fetch(T - 2, p2);
            const double* bwst = scr + (long long)(b1 - 1 - b0) * per_bw * N + (long long)3 * a.inner_it * N;
            pbx = bwst[0 * N + jj]; pbf = bwst[1 * N + jj]; pbn = bwst[2 * N + jj]; pbq = bwst[3 * N + jj];
            pbu = bwst[4 * N + jj];
        }
        for (int b = b1 - 1; b >= b0; b--) {
            if (own) {
                const double x = pbx, f = pbf, nu = pbn, q = pbq, u = pbu;
                if (b > b0) {
                    const double* bwst = scr + (long long)(b - 1 - b0) * per_bw * N + (long long)3 * a.inner_it * N;
                    pbx = bwst[0 * N + jj]; pbf = bwst[1 * N + jj]; pbn = bwst[2 * N + jj]; pbq = bwst[3 * N + jj];
                    pbu = bwst[4 * N + jj];
                }
                bw_adjoint(a, bw, b, x, f, nu, q, u, dbold, jj, lnr1, r);
            }
            // inner steps, last to first
            for (int i = a.inner_it - 1; i >= 0; i--) {
                double eE = 0.0, eI = 0.0, muE = 0.0, muI = 0.0, rE = 0.0;
                if (own) {
                    double v[5];
#pragma unroll
                    for (int q = 0; q < 5; q++) { v[q] = p1[q]; p1[q] = p2[q]; }
                    const int t = (b - b0) * a.inner_it + i;
                    if (t >= 2) fetch(t - 2, p2);
                    rww_adjoint_pre(c, v, G, wp, JN, wIE, sig, r, eE, eI, muE, muI, rE);
                }
                lds_barrier();
                if (act) ev[j] = eE * G * JN;
                lds_barrier();
                double coup = 0.0;
                if (a.sc_lds) {
#pragma unroll 8
                    for (int m = k0; m < k1; m++) coup += scs[jj * ldsc + m] * ev[m];
                } else {
#pragma unroll 8
                    for (int m = k0; m < k1; m++) coup += a.SC[(long long)jj * N + m] * ev[m];
                }
                if (ks > 1) {
                    if (!own) redk[(part - 1) * np + j] = coup;
                    lds_barrier();
                    if (own)
                        for (int p = 1; p < ks; p++) coup += redk[(p - 1) * np + j];
                }
                if (own) rww_adjoint_post(c, eE, eI, muE, muI, rE, wp, JN, wIE, coup, r);
            }
        }
    }
    store_gradients(a, sim, j, act, blockDim.x, r, red);
}

}  // namespace cubnm

using namespace cubnm;

namespace {

int fill_args(AdjArgs& A, RWW::C& c, BWConsts& bw, int n_sims, int nodes, int time_steps, int BOLD_TR, double dt,
              double bw_dt, int ck, int noise_time_steps, const double* SC, const double* G, const double* wp,
              const double* JN, const double* wIE, const double* sigma, const double* noise, const int* sh_nodes,
              const int* sh_ts) {
    if (n_sims <= 0 || nodes <= 0 || nodes > 1024) { set_last_error("need 0 < nodes <= 1024 and n_sims > 0"); return -1; }
    A = AdjArgs{};
    A.n_sims = n_sims; A.N = nodes;
    const double bw_dt_s = bw_dt / 1000.0;
    A.bw_it = (int)(((double)time_steps / 1000.0) / bw_dt_s);
    A.inner_it = (int)((bw_dt_s * 1000) / dt);
    A.TR_iters = (int)(((double)BOLD_TR / 1000.0) / bw_dt_s);
    A.bold_len = time_steps / BOLD_TR;
    A.noise_bw_it = (int)(((double)noise_time_steps / 1000.0) / bw_dt_s);
    if (A.bw_it <= 0 || A.inner_it <= 0 || A.TR_iters <= 0 || A.noise_bw_it <= 0 || ck <= 0) {
        set_last_error("invalid timing or checkpoint interval"); return -1;
    }
    A.ck = ck; A.n_ck = (A.bw_it + ck - 1) / ck;
    A.SC = SC; A.G = G; A.wp = wp; A.JN = JN; A.wIE = wIE; A.sigma = sigma;
    A.noise = noise; A.sh_nodes = sh_nodes; A.sh_ts = sh_ts;
    // k partitions: two when a second partition fits the 256-lane launch
    // bound and there are enough sources to split
    A.np = ((nodes + 63) / 64) * 64;
    A.ks = (A.np <= 128 && nodes >= 50) ? 2 : 1;
    RWW::init_constants(c, dt);
    bw = current_bw();
    bw.dt = bw_dt_s; bw.dt_itau = bw.dt * bw.itau;
    return 0;
}

int finish(hipError_t e) {
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) { set_last_error(std::string("HIP error: ") + hipGetErrorString(e)); return -1; }
    return 0;
}

}  // namespace

extern "C" {

int cubnm_rww_grad_sizes(int n_sims, int nodes, int time_steps, int BOLD_TR, double dt, double bw_dt, int ck,
                         long long* bold_doubles, long long* ckpt_doubles, long long* scratch_doubles) {
    AdjArgs A; RWW::C c; BWConsts bw;
    if (fill_args(A, c, bw, n_sims, nodes, time_steps, BOLD_TR, dt, bw_dt, ck, time_steps, nullptr, nullptr, nullptr,
                  nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) != 0)
        return -1;
    *bold_doubles = (long long)n_sims * A.bold_len * nodes;
    *ckpt_doubles = (long long)n_sims * A.n_ck * 6 * nodes;
    *scratch_doubles = (long long)n_sims * ck * (A.inner_it * 3 + 5) * nodes;
    return 0;
}

int cubnm_rww_grad_forward(int n_sims, int nodes, int time_steps, int BOLD_TR, double dt, double bw_dt, int ck,
                           int noise_time_steps, const double* d_SC, const double* d_G, const double* d_wp,
                           const double* d_JN, const double* d_wIE, const double* d_sigma, const double* d_noise,
                           const int* d_sh_nodes, const int* d_sh_ts, double* d_bold, double* d_ckpt) {
    AdjArgs A; RWW::C c; BWConsts bw;
    if (fill_args(A, c, bw, n_sims, nodes, time_steps, BOLD_TR, dt, bw_dt, ck, noise_time_steps, d_SC, d_G, d_wp, d_JN,
                  d_wIE, d_sigma, d_noise, d_sh_nodes, d_sh_ts) != 0)
        return -1;
    A.bold = d_bold; A.ckpt = d_ckpt;
    const size_t base = sizeof(double) * ((size_t)nodes + (size_t)(A.ks - 1) * A.np);
    const size_t with_sc = base + sizeof(double) * (size_t)nodes * (nodes + 1);
    A.sc_lds = with_sc <= 150 * 1024;
    const int threads = A.ks * A.np;
    if (threads <= 256)
        hipLaunchKernelGGL(rww_fwd_kernel<256>, dim3(n_sims), dim3(threads), A.sc_lds ? with_sc : base, 0, A, c, bw);
    else
        hipLaunchKernelGGL(rww_fwd_kernel<1024>, dim3(n_sims), dim3(threads), A.sc_lds ? with_sc : base, 0, A, c, bw);
    return finish(hipGetLastError());
}

int cubnm_rww_grad_backward(int n_sims, int nodes, int time_steps, int BOLD_TR, double dt, double bw_dt, int ck,
                            int noise_time_steps, const double* d_SC, const double* d_G, const double* d_wp,
                            const double* d_JN, const double* d_wIE, const double* d_sigma, const double* d_noise,
                            const int* d_sh_nodes, const int* d_sh_ts, const double* d_ckpt, const double* d_dbold,
                            double* d_scratch, double* d_gG, double* d_gwp, double* d_gJN) {
    AdjArgs A; RWW::C c; BWConsts bw;
    if (fill_args(A, c, bw, n_sims, nodes, time_steps, BOLD_TR, dt, bw_dt, ck, noise_time_steps, d_SC, d_G, d_wp, d_JN,
                  d_wIE, d_sigma, d_noise, d_sh_nodes, d_sh_ts) != 0)
        return -1;
    A.ckpt = const_cast<double*>(d_ckpt); A.dbold = d_dbold; A.scratch = d_scratch;
    A.gG = d_gG; A.gwp = d_gwp; A.gJN = d_gJN;
    const size_t base = sizeof(double) * (2 * (size_t)nodes + 16 + (size_t)(A.ks - 1) * A.np);
    const size_t with_sc = base + sizeof(double) * (size_t)nodes * (nodes + 1);
    A.sc_lds = with_sc <= 150 * 1024;
    const int threads = A.ks * A.np;
    if (threads <= 256)
        hipLaunchKernelGGL(rww_adj_kernel<256>, dim3(n_sims), dim3(threads), A.sc_lds ? with_sc : base, 0, A, c, bw);
    else
        hipLaunchKernelGGL(rww_adj_kernel<1024>, dim3(n_sims), dim3(threads), A.sc_lds ? with_sc : base, 0, A, c, bw);
    return finish(hipGetLastError());
}

}  // extern "C"
