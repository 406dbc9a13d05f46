// FP64 eigenvector refinement of the mixed-precision solver for large matrices.
//
// The FP32 reduction leaves eigenvector errors of order eps_32 ||A|| / gap, which
// for near-degenerate states at n_phi = 3000 are large enough to move plaquette
// phases across +-pi (the reference's FP32 builds lose Chern accuracy above
// n_phi ~ 1500, README.md:117). This file refines the FP32 eigenvectors X of each
// matrix with one step of the Ogita-Aishima iteration (X' = X + X E from
// S = X^H A X and G = X^H X), restricted to a window of 2 kRefW + 1 neighbouring
// states (eigenvalues ascending), after an exact FP64 Rayleigh-Ritz rotation inside
// every cluster of strongly coupled states:
//   ref_apply_kernel    Y = A X on the FP64 matrix cores (band form), X widened to FP64
//   ref_window_kernel   S_kj = x_k^H y_j and G_kj = x_k^H x_j for |k - j| <= kRefW
//   ref_cluster_kernel  runs of states whose first-order coupling
//                       |S_kj - l_j G_kj| / |l_j - l_k| exceeds kRefCoupling
//                       (l_j = S_jj / G_jj), in chunks of at most kRefMaxCluster:
//                       S_C q = l G_C q in FP64 (Cholesky, Jacobi), X_C <- X_C Q,
//                       Y_C <- Y_C Q
//   ref_window_kernel   again, for the rotated vectors
//   ref_update_kernel   X' = X + X E: E_kj = (S_kj - l_j G_kj) / (l_j - l_k) for k in
//                       another chunk, -G_kj / 2 within a chunk, E_jj = (1 - G_jj) / 2
// The plans then take the FP64 Rayleigh quotients and the links of X'.
// Layouts: X, Y, X' component-major per matrix ([i*n + j] = component i of state j,
// complex double); S, G: [kp][j][kRefWin] with entry d holding k = j + d - kRefW.
#include <mutex>

#include "common.hpp"
#include "refine.hpp"

#include <algorithm>

namespace rlm {

constexpr int kRefW = 16;
constexpr int kRefWin = 2 * kRefW + 1;
constexpr int kRefMaxCluster = 8;
constexpr double kRefCoupling = 1e-3;

int refine_window() { return kRefWin; }
int refine_max_cluster() { return kRefMaxCluster; }

// ----------------------------------------------------------------------------
// Y = A X for 16 states per wave (the band product of the Rayleigh kernel, stored),
// Xd = X widened to FP64. Four waves per workgroup, the workgroups of one matrix on
// one XCD.
template <typename TX>
__global__ __launch_bounds__(256) void ref_apply_kernel(const double *__restrict__ Hb, const TX *__restrict__ X,
                                                        int n, int bw, long nkp, int wg_per_kp,
                                                        double *__restrict__ Xd, double *__restrict__ Y) {
    using M = Mf<double>;
    using v4 = M::v4;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int li = lane & 15, lq = lane >> 4;
    const long q = blockIdx.x >> 3;
    const long kp = (q / wg_per_kp) * 8 + (blockIdx.x & 7);
    const int sb = (int)(q % wg_per_kp) * 4 + w;
    const int nsb = (n + 15) / 16;
    if (kp >= nkp || sb >= nsb) return;                      // no barriers in this kernel
    const int nb = band_nb(n, bw);
    const double *Hk = Hb + 2 * kp * (long)n * nb;
    const TX *Xk = X + 2 * kp * (long)n * n;
    double *Xdk = Xd + 2 * kp * (long)n * n;
    double *Yk = Y + 2 * kp * (long)n * n;
    const int j = sb * 16 + li;
    const bool jin = j < n;
    const bool dense = 16 + 2 * bw + 3 >= n;
    const int kw = dense ? n : 16 + 2 * bw;
    for (int r0 = 0; r0 < n; r0 += 16) {
        // three real products per complex one: q = (a_r + a_i) u_r, Re = q - a_i (u_r + u_i),
        // Im = q + a_r (u_i - u_r) (four waves of one state block each per SIMD hide the
        // gathers' latency better than fewer waves with several blocks: n_phi = 3000, 64
        // matrices, 33 ms against 52 ms with four blocks per wave)
        v4 yq = {0.0, 0.0, 0.0, 0.0}, yr = {0.0, 0.0, 0.0, 0.0}, yi = {0.0, 0.0, 0.0, 0.0};
        const int cs = dense ? 0 : r0 - bw;
        const int r = r0 + li;
        // four K-steps at a time: the band entry's offset and mode (0: outside the band,
        // 1: as stored, 2: conjugate; band_entry's formulas) and every load first, the
        // modes applied as each step's products issue (a gather per step, used at once,
        // left each step waiting for its loads)
        for (int k0 = 0; k0 < kw; k0 += 16) {
            cd araw[4];
            int md[4];
            cx<TX> u[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++) {
                const int kk = k0 + 4 * s4 + lq;
                int c = cs + kk;
                c = c < 0 ? c + n : (c >= n ? c - n : c);
                const bool kin = kk < kw;
                const bool up = r <= c;
                const int D = up ? c - r : r - c;
                const int b = (2 * bw + 1 >= n || D <= bw) ? D : (D >= n - bw ? D - n + 2 * bw + 1 : -1);
                const bool ok = kin && r < n && b >= 0;
                araw[s4] = ld(Hk, ok ? (long)(up ? r : c) * nb + b : 0);
                md[s4] = ok ? (up ? 2 : 1) : 0;
                u[s4] = (kin && jin) ? ldx(Xk, (long)c * n + j) : cx<TX>{(TX)0, (TX)0};
            }
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++) {
                const cd a = md[s4] == 0 ? cd{0.0, 0.0} : (md[s4] == 2 ? cconj(araw[s4]) : araw[s4]);
                const double ur = u[s4].x, ui = u[s4].y;
                yq = M::mma(a.x + a.y, ur, yq);
                yr = M::mma(-a.y, ur + ui, yr);
                yi = M::mma(a.x, ui - ur, yi);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int rr = r0 + M::row(lane, v);
            if (rr < n && jin) {
                st(Yk, (long)rr * n + j, cd{yq[v] + yr[v], yq[v] + yi[v]});
                const cx<TX> u = ldx(Xk, (long)rr * n + j);
                st(Xdk, (long)rr * n + j, cd{(double)u.x, (double)u.y});
            }
        }
    }
}

// ----------------------------------------------------------------------------
// S_kj = sum_i conj(X[i][k]) Y[i][j], G_kj = sum_i conj(X[i][k]) X[i][j] for the 16
// states j of a wave and the 48 states k = j0 - 16 .. j0 + 31 (three 16-wide tiles of
// FP64 MFMAs, K = components); the entries with |k - j| <= kRefW are stored.
__global__ __launch_bounds__(256) void ref_window_kernel(const double *__restrict__ Xd, const double *__restrict__ Y,
                                                         int n, long nkp, int wg_per_kp, double *__restrict__ Sw,
                                                         double *__restrict__ Gw) {
    using M = Mf<double>;
    using v4 = M::v4;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int li = lane & 15, lq = lane >> 4;
    const long q = blockIdx.x >> 3;
    const long kp = (q / wg_per_kp) * 8 + (blockIdx.x & 7);
    const int sb = (int)(q % wg_per_kp) * 4 + w;
    const int nsb = (n + 15) / 16;
    if (kp >= nkp || sb >= nsb) return;
    const double *Xk = Xd + 2 * kp * (long)n * n;
    const double *Yk = Y + 2 * kp * (long)n * n;
    const int j0 = sb * 16, jb = j0 + li;
    v4 sr[3], si[3], gr[3], gi[3];
#pragma unroll
    for (int t = 0; t < 3; t++) sr[t] = si[t] = gr[t] = gi[t] = v4{0.0, 0.0, 0.0, 0.0};
    for (int i0 = 0; i0 < n; i0 += 4) {
        const int i = i0 + lq;
        const bool iin = i < n;
        const cd yb = (iin && jb < n) ? ld(Yk, (long)i * n + jb) : cd{0.0, 0.0};
        const cd xb = (iin && jb < n) ? ld(Xk, (long)i * n + jb) : cd{0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const int k = j0 - 16 + 16 * t + li;
            const cd xa = (iin && k >= 0 && k < n) ? ld(Xk, (long)i * n + k) : cd{0.0, 0.0};
            // conj(xa) * yb and conj(xa) * xb
            sr[t] = M::mma(xa.x, yb.x, sr[t]);
            sr[t] = M::mma(xa.y, yb.y, sr[t]);
            si[t] = M::mma(xa.x, yb.y, si[t]);
            si[t] = M::mma(-xa.y, yb.x, si[t]);
            gr[t] = M::mma(xa.x, xb.x, gr[t]);
            gr[t] = M::mma(xa.y, xb.y, gr[t]);
            gi[t] = M::mma(xa.x, xb.y, gi[t]);
            gi[t] = M::mma(-xa.y, xb.x, gi[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < 3; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k = j0 - 16 + 16 * t + M::row(lane, v);
            const int d = k - jb;
            if (jb < n && k >= 0 && k < n && d >= -kRefW && d <= kRefW) {
                const long o = (kp * n + jb) * kRefWin + kRefW + d;
                st(Sw, o, cd{sr[t][v], si[t][v]});
                st(Gw, o, cd{gr[t][v], gi[t][v]});
            }
        }
}

// ----------------------------------------------------------------------------
// S q = l G q for a chunk of m <= kRefMaxCluster states, one thread, FP64:
// G = L L^H (Cholesky), M = L^-1 S L^-H, cyclic Jacobi M = V diag(l) V^H, Q = L^-H V
// with the eigenvalues ascending. Returns false if G is not positive definite.
__device__ bool small_geneig(int m, cd (&S)[kRefMaxCluster][kRefMaxCluster], cd (&G)[kRefMaxCluster][kRefMaxCluster],
                             cd (&Q)[kRefMaxCluster][kRefMaxCluster]) {
    constexpr int MC = kRefMaxCluster;
    cd L[MC][MC];
    for (int a = 0; a < m; a++)
        for (int b = 0; b < m; b++) L[a][b] = cd{0.0, 0.0};
    for (int c = 0; c < m; c++) {
        double dsum = G[c][c].x;
        for (int k = 0; k < c; k++) dsum -= cabs2(L[c][k]);
        if (!(dsum > 0.0)) return false;
        const double lcc = sqrt(dsum);
        L[c][c] = cd{lcc, 0.0};
        for (int r = c + 1; r < m; r++) {
            cd sacc = G[r][c];
            for (int k = 0; k < c; k++) sacc = csub(sacc, cmul(L[r][k], cconj(L[c][k])));
            L[r][c] = cscale(sacc, 1.0 / lcc);
        }
    }
    // T = L^-1 S (columns by forward substitution), M = (L^-1 T^H)^H
    cd T[MC][MC], Mt[MC][MC];
    for (int col = 0; col < m; col++)
        for (int r = 0; r < m; r++) {
            cd sacc = S[r][col];
            for (int k = 0; k < r; k++) sacc = csub(sacc, cmul(L[r][k], T[k][col]));
            T[r][col] = cscale(sacc, 1.0 / L[r][r].x);
        }
    for (int col = 0; col < m; col++)           // column col of L^-1 T^H = row col of M^H
        for (int r = 0; r < m; r++) {
            cd sacc = cconj(T[col][r]);
            for (int k = 0; k < r; k++) sacc = csub(sacc, cmul(L[r][k], Mt[k][col]));
            Mt[r][col] = cscale(sacc, 1.0 / L[r][r].x);
        }
    cd Mm[MC][MC], V[MC][MC];
    for (int a = 0; a < m; a++)
        for (int b = 0; b < m; b++) {
            Mm[a][b] = cconj(Mt[b][a]);
            V[a][b] = cd{a == b ? 1.0 : 0.0, 0.0};
        }
    for (int a = 0; a < m; a++) Mm[a][a].y = 0.0;
    for (int sweep = 0; sweep < 30; sweep++) {
        double off = 0.0, dia = 0.0;
        for (int a = 0; a < m; a++) {
            dia += Mm[a][a].x * Mm[a][a].x;
            for (int b = a + 1; b < m; b++) off += cabs2(Mm[a][b]);
        }
        if (off <= 1e-34 * dia || off == 0.0) break;
        for (int p = 0; p < m; p++)
            for (int qq = p + 1; qq < m; qq++) {
                const cd c = Mm[p][qq];
                const double r = sqrt(cabs2(c));
                if (r == 0.0) continue;
                const cd ph{c.x / r, -c.y / r};                   // e^{-i phi}
                const double tau = (Mm[qq][qq].x - Mm[p][p].x) / (2.0 * r);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
                // U = [[cs, sn], [-sn e^{-i phi}, cs e^{-i phi}]] on columns p, qq
                const cd upp{cs, 0.0}, upq{sn, 0.0}, uqp = cscale(ph, -sn), uqq = cscale(ph, cs);
                for (int k = 0; k < m; k++) {
                    const cd xp = Mm[k][p], xq = Mm[k][qq];
                    Mm[k][p] = cadd(cmul(xp, upp), cmul(xq, uqp));
                    Mm[k][qq] = cadd(cmul(xp, upq), cmul(xq, uqq));
                    const cd vp = V[k][p], vq = V[k][qq];
                    V[k][p] = cadd(cmul(vp, upp), cmul(vq, uqp));
                    V[k][qq] = cadd(cmul(vp, upq), cmul(vq, uqq));
                }
                for (int k = 0; k < m; k++) {
                    const cd xp = Mm[p][k], xq = Mm[qq][k];
                    Mm[p][k] = cadd(cmul(cconj(upp), xp), cmul(cconj(uqp), xq));
                    Mm[qq][k] = cadd(cmul(cconj(upq), xp), cmul(cconj(uqq), xq));
                }
                Mm[p][qq] = Mm[qq][p] = cd{0.0, 0.0};
                Mm[p][p].y = Mm[qq][qq].y = 0.0;
            }
    }
    // ascending order (selection sort of the columns)
    int ord[MC];
    for (int a = 0; a < m; a++) ord[a] = a;
    for (int a = 0; a < m; a++)
        for (int b = a + 1; b < m; b++)
            if (Mm[ord[b]][ord[b]].x < Mm[ord[a]][ord[a]].x) {
                const int tmp = ord[a];
                ord[a] = ord[b];
                ord[b] = tmp;
            }
    // Q = L^-H V (back substitution, columns in ascending order)
    for (int col = 0; col < m; col++)
        for (int r = m - 1; r >= 0; r--) {
            cd sacc = V[r][ord[col]];
            for (int k = r + 1; k < m; k++) sacc = csub(sacc, cmul(cconj(L[k][r]), Q[k][col]));
            Q[r][col] = cscale(sacc, 1.0 / L[r][r].x);
        }
    return true;
}

// One workgroup per matrix: chunks of coupled states, their Rayleigh-Ritz rotations
// (Qg[kp][j][b] = component b of the rotation column of state j) applied to X and Y,
// Cl[kp][j] = first state of j's chunk (-1 outside chunks). nflag[kp] counts chunks
// cut at kRefMaxCluster and failed Cholesky factorisations (diagnostic).
__global__ __launch_bounds__(256) void ref_cluster_kernel(double *__restrict__ Xd, double *__restrict__ Y, int n,
                                                          const double *__restrict__ Sw,
                                                          const double *__restrict__ Gw, double *__restrict__ Qg,
                                                          int *__restrict__ Cl, int *__restrict__ nflag) {
    extern __shared__ double lds_ref[];
    double *lam = lds_ref;                                   // n
    int *bound = reinterpret_cast<int *>(lam + n);           // n: states b and b+1 coupled
    int *clist = bound + n;                                  // (first state, size) pairs, <= n/2
    __shared__ int ncl;
    const long kp = blockIdx.x;
    const int t = threadIdx.x;
    const double *Sk = Sw + 2 * kp * (long)n * kRefWin;
    const double *Gk = Gw + 2 * kp * (long)n * kRefWin;
    int *Ck = Cl + kp * (long)n;
    if (t == 0) {
        ncl = 0;
        nflag[kp] = 0;
    }
    for (int j = t; j < n; j += 256) {
        lam[j] = Sk[2 * ((long)j * kRefWin + kRefW)] / Gk[2 * ((long)j * kRefWin + kRefW)];
        bound[j] = 0;
        Ck[j] = -1;
    }
    __syncthreads();
    for (int j = t; j < n; j += 256)
        for (int d = 1; d <= kRefW && j - d >= 0; d++) {
            const int k = j - d;
            const long o = (long)j * kRefWin + kRefW - d;   // (k, j)
            const cd s = ld(Sk, o), g = ld(Gk, o);
            const cd num = csub(s, cscale(g, lam[j]));
            if (sqrt(cabs2(num)) > kRefCoupling * fabs(lam[j] - lam[k]))
                for (int b = k; b < j; b++) bound[b] = 1;   // same value from every writer
        }
    __syncthreads();
    for (int s0 = t; s0 < n - 1; s0 += 256) {
        if (!bound[s0] || (s0 > 0 && bound[s0 - 1])) continue;
        int e = s0;
        while (e < n - 1 && bound[e]) e++;                  // run [s0, e]
        for (int c0 = s0; c0 <= e; c0 += kRefMaxCluster) {
            const int m = min(kRefMaxCluster, e - c0 + 1);
            if (c0 > s0) atomicAdd(nflag + kp, 1);
            if (m < 2) continue;
            cd S[kRefMaxCluster][kRefMaxCluster], G[kRefMaxCluster][kRefMaxCluster], Q[kRefMaxCluster][kRefMaxCluster];
            for (int a = 0; a < m; a++)
                for (int b = 0; b < m; b++) {
                    // entry (k = c0 + a, j = c0 + b) sits in row j at d = a - b
                    const long o = (long)(c0 + b) * kRefWin + kRefW + (a - b);
                    S[a][b] = ld(Sk, o);
                    G[a][b] = ld(Gk, o);
                }
            if (!small_geneig(m, S, G, Q)) {
                atomicAdd(nflag + kp, 1);
                continue;
            }
            for (int a = 0; a < m; a++) {
                Ck[c0 + a] = c0;
                for (int b = 0; b < kRefMaxCluster; b++)
                    st(Qg, (kp * n + c0 + a) * kRefMaxCluster + b, b < m ? Q[b][a] : cd{0.0, 0.0});
            }
            const int slot = atomicAdd(&ncl, 1);
            clist[2 * slot] = c0;
            clist[2 * slot + 1] = m;
        }
    }
    __syncthreads();
    // X_C <- X_C Q, Y_C <- Y_C Q, row by row
    const int nc = ncl;
    double *Xk = Xd + 2 * kp * (long)n * n;
    double *Yk = Y + 2 * kp * (long)n * n;
    for (int c = 0; c < nc; c++) {
        const int c0 = clist[2 * c], m = clist[2 * c + 1];
        for (int i = t; i < 2 * n; i += 256) {
            double *Bk = i < n ? Xk : Yk;
            const long row = (long)(i < n ? i : i - n) * n;
            cd old[kRefMaxCluster];
            for (int b = 0; b < m; b++) old[b] = ld(Bk, row + c0 + b);
            for (int a = 0; a < m; a++) {
                cd acc{0.0, 0.0};
                for (int b = 0; b < m; b++) cfma(acc, old[b], ld(Qg, (kp * n + c0 + a) * kRefMaxCluster + b));
                st(Bk, row + c0 + a, acc);
            }
        }
    }
}

// ----------------------------------------------------------------------------
// X' = X + X E for 64 states per workgroup (E in LDS), threads over (row, state).
__global__ __launch_bounds__(256) void ref_update_kernel(const double *__restrict__ Xd, const double *__restrict__ Sw,
                                                         const double *__restrict__ Gw, const int *__restrict__ Cl,
                                                         int n, int wpk, double *__restrict__ Xout) {
    __shared__ cd E[64][kRefWin];
    __shared__ cd xrow[4][64 + 2 * kRefW];
    const long kp = blockIdx.x / wpk;
    const int j0 = (int)(blockIdx.x % wpk) * 64;
    const int t = threadIdx.x;
    const double *Sk = Sw + 2 * kp * (long)n * kRefWin;
    const double *Gk = Gw + 2 * kp * (long)n * kRefWin;
    const int *Ck = Cl + kp * (long)n;
    for (int e = t; e < 64 * kRefWin; e += 256) {
        const int jj = e / kRefWin, dd = e % kRefWin;
        const int j = j0 + jj, k = j + dd - kRefW;
        cd v{0.0, 0.0};
        if (j < n && k >= 0 && k < n) {
            const long oj = (long)j * kRefWin;
            const cd s = ld(Sk, oj + dd), g = ld(Gk, oj + dd);
            const double gjj = Gk[2 * (oj + kRefW)];
            const double lj = Sk[2 * (oj + kRefW)] / gjj;
            if (k == j) {
                v = cd{0.5 * (1.0 - gjj), 0.0};
            } else if (Ck[j] >= 0 && Ck[j] == Ck[k]) {
                v = cscale(g, -0.5);
            } else {
                const long ok = (long)k * kRefWin + kRefW;
                const double lk = Sk[2 * ok] / Gk[2 * ok];
                v = cscale(csub(s, cscale(g, lj)), 1.0 / (lj - lk));
            }
        }
        E[jj][dd] = v;
    }
    __syncthreads();
    // each wave takes every fourth row: the row's window X[i][j0 - kRefW .. j0 + 63 + kRefW]
    // is staged in the wave's LDS buffer (two coalesced loads per lane instead of kRefWin
    // scattered ones), then every lane forms its state's entry from it
    const int jj = t & 63, j = j0 + jj, wv = t >> 6;
    cd *xb = xrow[wv];
    const double *Xk = Xd + 2 * kp * (long)n * n;
    double *Ok = Xout + 2 * kp * (long)n * n;
    for (int i = wv; i < n; i += 4) {
        const long row = (long)i * n;
        for (int e = jj; e < 64 + 2 * kRefW; e += 64) {
            const int k = j0 - kRefW + e;
            xb[e] = (k >= 0 && k < n) ? ld(Xk, row + k) : cd{0.0, 0.0};
        }
        __builtin_amdgcn_wave_barrier();
        if (j < n) {
            cd acc = xb[jj + kRefW];
            for (int dd = 0; dd < kRefWin; dd++) cfma(acc, xb[jj + dd], E[jj][dd]);
            st(Ok, row + j, acc);
        }
        __builtin_amdgcn_wave_barrier();                     // the buffer is rewritten by the next row
    }
}

// ----------------------------------------------------------------------------
// Cluster Rayleigh-Ritz for the mixed solver without the full refinement: after the
// FP64 Rayleigh quotients, runs of neighbouring states whose eigenvalue gaps are below
// kPairTau times the spectrum's extent (near-degenerate states, whose FP32 vectors mix
// and whose Rayleigh quotients then err by |coupling|^2 / gap) are collected, and each
// run gets an exact FP64 Rayleigh-Ritz: y_a = A u_a from the band form, S_ab = u_a^H
// y_b, G_ab = u_a^H u_b, S q = l G q, eigenvalues into W and the FP32 vectors rotated by
// Q. Pairs (the common case) take a register-resident 2 x 2 solver, runs of 3 to
// kPairMax states an LDS-staged general one; longer runs are split into balanced chunks
// of at most kPairMax (never a single state) and counted. The threshold: the FP32
// vectors' couplings |u_k^H A u_j| measure ~5e-7 at n_phi = 256, so a gap g leaves
// |coupling|^2 / g in the Rayleigh quotient, below 1e-10 of the spectrum's extent for g
// above ~1e-4 of it.
constexpr double kPairTau = 1e-4;
constexpr int kPairMax = kRefMaxCluster;       // 8
constexpr int kPairMaxN = 512;                 // n of the unrefined mixed path (LDS stage)

// count[0] = runs (chunks) appended, count[1] = runs longer than kPairMax (split),
// count[2] = chunks whose Gram matrix had no Cholesky factor (left unrotated)
__global__ void pair_zero_kernel(int *count) {
    if (threadIdx.x < kPairCounters) count[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void pair_flag_kernel(const double *__restrict__ W, int n, long nkp, int cap,
                                                        int *__restrict__ list, int *__restrict__ count) {
    const long kp = blockIdx.y;
    if (kp >= nkp) return;
    const double *w = W + kp * n;
    const double tau = kPairTau * (fabs(w[0]) + fabs(w[n - 1]));
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n - 1; j += gridDim.x * blockDim.x) {
        const bool here = w[j + 1] - w[j] < tau;
        const bool before = j > 0 && w[j] - w[j - 1] < tau;
        if (!here || before) continue;                        // not the first state of a run
        int e = j + 1;                                        // run [j, e]
        while (e < n - 1 && w[e + 1] - w[e] < tau) e++;
        const int len = e - j + 1;
        const int nch = (len + kPairMax - 1) / kPairMax;      // balanced chunks, each >= 2 states
        if (nch > 1) atomicAdd(count + 1, 1);
        const int base = len / nch, rem = len % nch;
        for (int q = 0, c0 = j; q < nch; q++) {
            const int m = base + (q < rem ? 1 : 0);
            const int slot = atomicAdd(count, 1);
            if (slot < cap) {
                list[3 * slot] = (int)kp;
                list[3 * slot + 1] = c0;
                list[3 * slot + 2] = m;
            }
            c0 += m;
        }
    }
}

// small_geneig for m = 2 in registers: the same Cholesky G = L L^H, M = L^-1 S L^-H,
// one Jacobi rotation (exact for 2 x 2), ascending order and Q = L^-H V.
__host__ __device__ inline bool small_geneig2(const cd (&S)[2][2], const cd (&G)[2][2], cd (&Q)[2][2]) {
    const double g11 = G[0][0].x;
    if (!(g11 > 0.0)) return false;
    const double l11 = sqrt(g11);
    const cd l21 = cscale(G[1][0], 1.0 / l11);
    const double d2 = G[1][1].x - cabs2(l21);
    if (!(d2 > 0.0)) return false;
    const double l22 = sqrt(d2);
    // X = L^-1 S, M = X L^-H
    cd X[2][2];
    for (int c = 0; c < 2; c++) {
        X[0][c] = cscale(S[0][c], 1.0 / l11);
        X[1][c] = cscale(csub(S[1][c], cmul(l21, X[0][c])), 1.0 / l22);
    }
    const double a = X[0][0].x / l11;
    const double d = (csub(X[1][1], cmul(X[1][0], cscale(cconj(l21), 1.0 / l11)))).x / l22;
    const cd b = cscale(csub(X[0][1], cmul(X[0][0], cscale(cconj(l21), 1.0 / l11))), 1.0 / l22);
    // rotation [[cs, sn], [-sn e^{-i phi}, cs e^{-i phi}]], b = |b| e^{i phi}: diagonal a - t|b|, d + t|b|
    const double r = sqrt(cabs2(b));
    cd V[2][2] = {{cd{1.0, 0.0}, cd{0.0, 0.0}}, {cd{0.0, 0.0}, cd{1.0, 0.0}}};
    double l0 = a, l1 = d;
    if (r > 0.0) {
        const cd ph{b.x / r, -b.y / r};
        const double tau = (d - a) / (2.0 * r);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
        V[0][0] = cd{cs, 0.0};
        V[0][1] = cd{sn, 0.0};
        V[1][0] = cscale(ph, -sn);
        V[1][1] = cscale(ph, cs);
        l0 = a - t * r;
        l1 = d + t * r;
    }
    const int o0 = l1 < l0 ? 1 : 0, o1 = 1 - o0;
    for (int col = 0; col < 2; col++) {
        const int src = col == 0 ? o0 : o1;
        const cd v0 = src ? V[0][1] : V[0][0], v1 = src ? V[1][1] : V[1][0];
        Q[1][col] = cscale(v1, 1.0 / l22);
        Q[0][col] = cscale(csub(v0, cmul(cconj(l21), Q[1][col])), 1.0 / l11);
    }
    return true;
}

// One pair (states c0, c0 + 1) at kp: one wave, lanes over rows, P and G in registers.
// S = U^H A U from the upper half of the band, A = B + B^H with B(r, c) = A(r, c) at
// circular offsets c - r = 1..bw and B(r, r) = A(r, r) / 2 (the upper triangle when the
// band covers every column): P = U^H B U, S = P + P^H.
__device__ void pair_rr_two(const double *Ak, const double *Hk, float *Uk, double *Wk, int n, int bw, int nb,
                            bool dense, int c0, int lane, cx<float> (*ust)[kPairMaxN], cd (*Qs)[2], int *okflag,
                            int *nfail) {
    for (int e = lane; e < 2 * n; e += 64) {
        const int a = e / n, r = e - a * n;
        ust[a][r] = ldx(Uk, (long)r * n + c0 + a);
    }
    __syncthreads();
    cd P[2][2], G[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) P[a][b] = G[a][b] = cd{0.0, 0.0};
    for (int r = lane; r < n; r += 64) {
        cd u[2], y[2];
#pragma unroll
        for (int a = 0; a < 2; a++) {
            u[a] = cd{ust[a][r].x, ust[a][r].y};
            y[a] = cd{0.0, 0.0};
        }
        const int nc = dense ? n - r : bw + 1;
        // band form: entries (r, r + k) with r + k < n are conj(H[r][k]), a contiguous run
        // of this row, read without band_entry's branches so that the loads batch; the same
        // terms in the same order as band_entry gives them
        const int k1 = Hk ? min(nc, n - r) : 0;
        const cd *hr = Hk ? reinterpret_cast<const cd *>(Hk) + (long)r * nb : nullptr;
        int k = 0;
        if (k1 > 0) {
            const cd arc{0.5 * hr[0].x, 0.0};
#pragma unroll
            for (int a = 0; a < 2; a++) cfma(y[a], arc, cd{ust[a][r].x, ust[a][r].y});
            k = 1;
        }
#pragma unroll 4
        for (; k < k1; k++) {
            const cd arc = cconj(hr[k]);
#pragma unroll
            for (int a = 0; a < 2; a++) cfma(y[a], arc, cd{ust[a][r + k].x, ust[a][r + k].y});
        }
        for (; k < nc; k++) {
            const int c = r + k >= n ? r + k - n : r + k;
            cd arc = band_entry(Ak, Hk, n, bw, nb, r, c);
            if (k == 0) arc = cd{0.5 * arc.x, 0.0};
#pragma unroll
            for (int a = 0; a < 2; a++) cfma(y[a], arc, cd{ust[a][c].x, ust[a][c].y});
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                cfmac(P[a][b], u[a], y[b]);
                cfmac(G[a][b], u[a], u[b]);
            }
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            P[a][b] = cd{wave_sum(P[a][b].x), wave_sum(P[a][b].y)};
            G[a][b] = cd{wave_sum(G[a][b].x), wave_sum(G[a][b].y)};
        }
    if (lane == 0) {
        cd Sq[2][2], Gq[2][2], Q[2][2];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                Sq[a][b] = cadd(P[a][b], cconj(P[b][a]));
                Gq[a][b] = cscale(cadd(G[a][b], cconj(G[b][a])), 0.5);
            }
        const bool ok = small_geneig2(Sq, Gq, Q);
        if (ok) {
#pragma unroll
            for (int a = 0; a < 2; a++) {
                // l_a = q_a^H S q_a (Q^H G Q = I)
                cd acc{0.0, 0.0};
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    cd t{0.0, 0.0};
#pragma unroll
                    for (int c = 0; c < 2; c++) cfma(t, Sq[b][c], Q[c][a]);
                    cfmac(acc, Q[b][a], t);
                }
                Wk[c0 + a] = acc.x;
#pragma unroll
                for (int b = 0; b < 2; b++) Qs[b][a] = Q[b][a];
            }
        } else {
            atomicAdd(nfail, 1);
        }
        *okflag = ok ? 1 : 0;
    }
    __syncthreads();
    // u_a <- sum_b u_b Q[b][a] (skipped when G had no Cholesky factor)
    if (*okflag) {
        for (int r = lane; r < n; r += 64) {
            const cd u0{ust[0][r].x, ust[0][r].y}, u1{ust[1][r].x, ust[1][r].y};
#pragma unroll
            for (int a = 0; a < 2; a++) {
                cd acc{0.0, 0.0};
                cfma(acc, u0, Qs[0][a]);
                cfma(acc, u1, Qs[1][a]);
                stx(Uk, (long)r * n + c0 + a, cx<float>{(float)acc.x, (float)acc.y});
            }
        }
    }
    __syncthreads();
}

// One wave per pair (persistent over the list; chunks of 3 or more states are left to
// pair_rr_long_kernel).
__global__ __launch_bounds__(64) void pair_rr_kernel(const double *__restrict__ A, const double *__restrict__ Hb,
                                                     float *__restrict__ U, int n, int bw, int cap,
                                                     const int *__restrict__ list, int *__restrict__ count,
                                                     double *__restrict__ W) {
    __shared__ cd Qs[2][2];
    __shared__ int okflag;
    __shared__ cx<float> ust[2][kPairMaxN];
    const int lane = threadIdx.x;
    const int total = min(count[0], cap);
    const int nb = band_nb(n, bw);
    const bool dense = 2 * bw + 1 >= n;
    for (int r_ = blockIdx.x; r_ < total; r_ += gridDim.x) {
        if (list[3 * r_ + 2] != 2) continue;
        const long kp = list[3 * r_];
        const int c0 = list[3 * r_ + 1];
        const double *Hk = Hb ? Hb + 2 * kp * (long)n * nb : nullptr;
        const double *Ak = Hb ? nullptr : A + 2 * kp * (long)n * n;
        pair_rr_two(Ak, Hk, U + 2 * kp * (long)n * n, W + kp * n, n, bw, nb, dense, c0, lane, ust, Qs, &okflag,
                    count + 2);
    }
}

// Chunks of 3..kPairMax states: one wave per chunk; the chunk's vectors (FP32) and
// y_a = B u_a (FP64) staged in LDS (dynamic: kPairMax * kPairMaxN * 24 bytes), lane
// (a, b) forms P_ab and G_ab, lane 0 solves S q = l G q (small_geneig), all lanes rotate.
__global__ __launch_bounds__(64) void pair_rr_long_kernel(const double *__restrict__ A, const double *__restrict__ Hb,
                                                          float *__restrict__ U, int n, int bw, int cap,
                                                          const int *__restrict__ list, int *__restrict__ count,
                                                          double *__restrict__ W) {
    constexpr int MC = kPairMax;
    extern __shared__ unsigned char lds_raw[];
    cd(*ys)[kPairMaxN] = reinterpret_cast<cd(*)[kPairMaxN]>(lds_raw);
    cx<float>(*ust)[kPairMaxN] = reinterpret_cast<cx<float>(*)[kPairMaxN]>(lds_raw + sizeof(cd) * MC * kPairMaxN);
    __shared__ cd Ps[MC][MC], Gs[MC][MC], Qs[MC][MC];
    __shared__ int okflag;
    const int lane = threadIdx.x;
    const int total = min(count[0], cap);
    const int nb = band_nb(n, bw);
    const bool dense = 2 * bw + 1 >= n;
    for (int r_ = blockIdx.x; r_ < total; r_ += gridDim.x) {
        const int m = list[3 * r_ + 2];
        if (m <= 2) continue;
        const long kp = list[3 * r_];
        const int c0 = list[3 * r_ + 1];
        const double *Hk = Hb ? Hb + 2 * kp * (long)n * nb : nullptr;
        const double *Ak = Hb ? nullptr : A + 2 * kp * (long)n * n;
        float *Uk = U + 2 * kp * (long)n * n;
        double *Wk = W + kp * n;
        for (int e = lane; e < MC * n; e += 64) {
            const int a = e / n, r = e - a * n;
            ust[a][r] = a < m ? ldx(Uk, (long)r * n + c0 + a) : cx<float>{0.f, 0.f};
        }
        __syncthreads();
        for (int r = lane; r < n; r += 64) {
            cd y[MC];
#pragma unroll
            for (int a = 0; a < MC; a++) y[a] = cd{0.0, 0.0};
            const int nc = dense ? n - r : bw + 1;
            for (int k = 0; k < nc; k++) {
                const int c = r + k >= n ? r + k - n : r + k;
                cd arc = band_entry(Ak, Hk, n, bw, nb, r, c);
                if (k == 0) arc = cd{0.5 * arc.x, 0.0};
#pragma unroll
                for (int a = 0; a < MC; a++) cfma(y[a], arc, cd{ust[a][c].x, ust[a][c].y});
            }
#pragma unroll
            for (int a = 0; a < MC; a++) ys[a][r] = y[a];
        }
        __syncthreads();
        {
            const int a = lane >> 3, b = lane & 7;
            cd p{0.0, 0.0}, gg{0.0, 0.0};
            if (a < m && b < m)
                for (int r = 0; r < n; r++) {
                    const cd ua{ust[a][r].x, ust[a][r].y}, ub{ust[b][r].x, ust[b][r].y};
                    cfmac(p, ua, ys[b][r]);
                    cfmac(gg, ua, ub);
                }
            Ps[a][b] = p;
            Gs[a][b] = gg;
        }
        __syncthreads();
        if (lane == 0) {
            cd Sq[MC][MC], Gq[MC][MC], Q[MC][MC];
            for (int a = 0; a < MC; a++)
                for (int b = 0; b < MC; b++) {
                    const bool in = a < m && b < m;
                    Sq[a][b] = in ? cadd(Ps[a][b], cconj(Ps[b][a])) : cd{0.0, 0.0};
                    Gq[a][b] = in ? cscale(cadd(Gs[a][b], cconj(Gs[b][a])), 0.5) : cd{0.0, 0.0};
                }
            const bool ok = small_geneig(m, Sq, Gq, Q);
            if (ok) {
                for (int a = 0; a < m; a++) {
                    cd acc{0.0, 0.0};
                    for (int b = 0; b < m; b++) {
                        cd t{0.0, 0.0};
                        for (int c = 0; c < m; c++) cfma(t, Sq[b][c], Q[c][a]);
                        cfmac(acc, Q[b][a], t);
                    }
                    Wk[c0 + a] = acc.x;
                }
                for (int a = 0; a < MC; a++)
                    for (int b = 0; b < MC; b++) Qs[b][a] = (a < m && b < m) ? Q[b][a] : cd{0.0, 0.0};
            } else {
                atomicAdd(count + 2, 1);
            }
            okflag = ok ? 1 : 0;
        }
        __syncthreads();
        if (okflag) {
            for (int r = lane; r < n; r += 64) {
                cd u[MC];
#pragma unroll
                for (int b = 0; b < MC; b++) u[b] = cd{ust[b][r].x, ust[b][r].y};
                for (int a = 0; a < m; a++) {
                    cd acc{0.0, 0.0};
#pragma unroll
                    for (int b = 0; b < MC; b++) cfma(acc, u[b], Qs[b][a]);
                    stx(Uk, (long)r * n + c0 + a, cx<float>{(float)acc.x, (float)acc.y});
                }
            }
        }
        __syncthreads();
    }
}

int pair_list_capacity(int n, long nkp) { return (int)std::min<long>(nkp * (long)(n / 2), 1L << 26); }

int launch_pair_rr(const double *A, const double *Hb, float *U32, int n, int bw, long nkp, double *W, int *list,
                   int cap, int *count, hipStream_t s) {
    if (nkp <= 0 || n < 2) return 0;
    if (n > kPairMaxN) return -22;
    hipLaunchKernelGGL(pair_zero_kernel, dim3(1), dim3(64), 0, s, count);
    hipLaunchKernelGGL(pair_flag_kernel, dim3((n + 255) / 256, (unsigned)nkp), dim3(256), 0, s, W, n, nkp, cap, list,
                       count);
    hipLaunchKernelGGL(pair_rr_kernel, dim3(4096), dim3(64), 0, s, A, Hb, U32, n, bw, cap, list, count, W);
    const size_t lds = (sizeof(cd) + sizeof(cx<float>)) * kPairMax * kPairMaxN;
    static std::once_flag once;
    static hipError_t attr = hipSuccess;
    std::call_once(once, [&] {
        attr = hipFuncSetAttribute((const void *)pair_rr_long_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds);
    });
    RLM_HIP_CHECK(attr);
    hipLaunchKernelGGL(pair_rr_long_kernel, dim3(256), dim3(64), lds, s, A, Hb, U32, n, bw, cap, list, count, W);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

// ----------------------------------------------------------------------------
int launch_refine(const double *Hb, const float *U32, int n, int bw, long nkp, double *Xd, double *Y, double *Sw,
                  double *Gw, double *Qg, int *Cl, int *nflag, hipStream_t s) {
    if (nkp <= 0) return 0;
    const int nsb = (n + 15) / 16, wg = (nsb + 3) / 4;
    const long groups = (nkp + 7) / 8;
    const dim3 gw((unsigned)(groups * 8 * wg));
    hipLaunchKernelGGL(ref_apply_kernel<float>, gw, dim3(256), 0, s, Hb, U32, n, bw, nkp, wg, Xd, Y);
    hipLaunchKernelGGL(ref_window_kernel, gw, dim3(256), 0, s, Xd, Y, n, nkp, wg, Sw, Gw);
    const size_t lds = sizeof(double) * n + sizeof(int) * (n + 2 * (n / 2 + 1));
    if (lds > 64 * 1024)
        RLM_HIP_CHECK(
            hipFuncSetAttribute((const void *)ref_cluster_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(ref_cluster_kernel, dim3((unsigned)nkp), dim3(256), lds, s, Xd, Y, n, Sw, Gw, Qg, Cl, nflag);
    hipLaunchKernelGGL(ref_window_kernel, gw, dim3(256), 0, s, Xd, Y, n, nkp, wg, Sw, Gw);
    const int wpk = (n + 63) / 64;
    hipLaunchKernelGGL(ref_update_kernel, dim3((unsigned)(nkp * wpk)), dim3(256), 0, s, Xd, Sw, Gw, Cl, n, wpk, Y);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace rlm
