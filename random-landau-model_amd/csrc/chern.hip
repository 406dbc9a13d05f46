// Chern stage and mesh-averaged energies (FP64).
//
// Reference: diag_wrappers/chern.cpp:37-75 (peer_cal_Chern: per plaquette and
// state, prod of the 4 link overlaps zdotc(u_l, u_r) around
// (t1-1,m)->(t1,m)->(t1,m+1)->(t1-1,m+1), chern_theta = -arg(prod)/2pi),
// rlm/main_rlm_cpu.cpp:155-158,187-190,230-239 (energy sums over the
// (g+1)^2 mesh, Chern sums line by line, +1/n_phi) and
// hlm/main_hlm_cpu.cpp:111-118,143-150,184-191 (theta_2 = g copied from
// theta_2 = 0, energies sum theta_2 < g plus theta_2 = 0, Chern = -sum).
//
// Instead of walking 4 links per plaquette, every link is computed once
// (two per twist point), then a per-state kernel walks the plaquettes in the
// reference's order and accumulates in FP64.
//
// Eigenvectors are component-major: U[kp][i*ns + j] is component i of state j,
// so one thread per state reads coalesced rows.
#include "common.hpp"
#include "chern.hpp"

namespace rlm {

// links: grid (nkp), block 256 (threads over states).
// Twist point kp = (s*lines + t1)*L2n + t2 with L2n stored theta_2 per line
// (g+1 for rlm, g for hlm whose theta_2 = g is theta_2 = 0) and nlink2 links
// along theta_2 per line (g in both models).
// L1[kp][j] = <u(t1,t2)|u(t1+1,t2)> (t1 < lines-1); L2[kp][j] = <u(t1,t2)|u(t1,t2+1)>.
// TU = double (FP64 path) or float (mixed path: products of two floats are exact in
// double, so the FP64 accumulation is a compensated dot product of the FP32 vectors).
template <typename TU>
__global__ __launch_bounds__(256) void links_kernel(const TU *__restrict__ U, long strideU, int dim_vec, int ns,
                                                    int lines, int L2n, int nlink2, long kp0,
                                                    double *__restrict__ Lk1, double *__restrict__ Lk2) {
    const long kp = kp0 + blockIdx.x;
    const int t2 = (int)(kp % L2n);
    const int t1 = (int)((kp / L2n) % lines);
    const long base_line = kp - t2;         // kp of (t1, 0)
    const bool has1 = t1 < lines - 1;
    const bool has2 = t2 < nlink2;
    const long kp_1 = kp + L2n;             // (t1+1, t2)
    const long kp_2 = base_line + ((t2 + 1) % L2n);  // (t1, t2+1), wraps to 0 for hlm
    const TU *u0 = U + 2 * (kp - kp0) * strideU;
    const TU *u1 = U + 2 * (kp_1 - kp0) * strideU;
    const TU *u2 = U + 2 * (kp_2 - kp0) * strideU;
    auto up = [](cx<TU> v) { return cd{(double)v.x, (double)v.y}; };
    for (int j = blockIdx.y * blockDim.x + threadIdx.x; j < ns; j += gridDim.y * blockDim.x) {
        cd a1{0.0, 0.0}, a2{0.0, 0.0};
        for (int i = 0; i < dim_vec; i++) {
            const cd x = up(ldx(u0, (long)i * ns + j));
            if (has1) cfmac(a1, x, up(ldx(u1, (long)i * ns + j)));
            if (has2) cfmac(a2, x, up(ldx(u2, (long)i * ns + j)));
        }
        st(Lk1, kp * ns + j, a1);
        st(Lk2, kp * ns + j, a2);
    }
}

// rlm links with the twist-boundary points derived instead of solved. U holds the
// solved points (theta_1, theta_2 < g), kp_s = (s*g + t1)*g + t2; the others follow
// exactly from the model's tables (the matrix at theta_2 = 2 pi is the theta_2 = 0
// matrix with rows and columns shifted by one, the one at theta_1 = 2 pi is
// diag(e^{i 2 pi j/N}) A diag(e^{-i 2 pi j/N})):
//   u(t1, g)[i] = u(t1, 0)[(i+1) % N],  u(g, t2)[i] = e^{i 2 pi i/N} u(0, t2)[i].
// Links are written for the full (g+1)^2 mesh, kp = (s*(g+1) + t1)*(g+1) + t2, as
// links_kernel does; grid (8 * ceil(nkp / 8), state chunks of 256), block 256, ns
// complex doubles of dynamic LDS. Workgroups go to the 8 XCDs round-robin; with FP64
// vectors XCD x (blockIdx.x % 8) takes the contiguous run of points x * per .. and
// walks it in order, so that the neighbours (t1, t2 + 1) and (t1 + 1, t2) a point reads
// are read by the same XCD's workgroups at about the same time, from its L2 (n_phi =
// 256, 34848 points: 13.1 -> 11.6 ms); FP32 vectors keep the dispatch order, which
// measures faster for them (7.2 against 7.8 ms).
template <typename TU>
__global__ __launch_bounds__(256) void links_derived_kernel(const TU *__restrict__ U, long strideU, int ns, int g,
                                                            long nkp, double *__restrict__ Lk1,
                                                            double *__restrict__ Lk2) {
    extern __shared__ cd phs[];                              // e^{i 2 pi i/N}
    const long per = (nkp + 7) / 8;
    const long kp = sizeof(TU) == sizeof(double) ? (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3) : (long)blockIdx.x;
    if (kp >= nkp) return;                                   // whole workgroup, before any barrier
    const int g1 = g + 1;
    const long s = kp / ((long)g1 * g1);
    const int rem = (int)(kp % ((long)g1 * g1));
    const int t1 = rem / g1, t2 = rem % g1;
    const double TWO_PI = 6.283185307179586232;
    if (t1 >= g - 1) {                                       // blocks that touch theta_1 = 2 pi
        for (int i = threadIdx.x; i < ns; i += blockDim.x) {
            double sn, cs;
            sincos(TWO_PI * (double)i / (double)ns, &sn, &cs);
            phs[i] = cd{cs, sn};
        }
    }
    __syncthreads();
    // point (a, b) -> solved base, component shift, component phase
    auto src = [&](int a, int b, const TU *&base, int &shift, bool &phase) {
        shift = (b == g) ? 1 : 0;
        phase = (a == g);
        const int a0 = (a == g) ? 0 : a, b0 = (b == g) ? 0 : b;
        base = U + 2 * (((s * g + a0) * g + b0) * strideU);
    };
    const TU *u0, *u1, *u2;
    int sh0, sh1, sh2;
    bool ph0, ph1, ph2;
    const bool has1 = t1 < g, has2 = t2 < g;
    src(t1, t2, u0, sh0, ph0);
    src(has1 ? t1 + 1 : t1, t2, u1, sh1, ph1);
    src(t1, has2 ? t2 + 1 : t2, u2, sh2, ph2);
    auto comp = [&](const TU *u, int sh, bool ph, int i, int j) -> cd {
        int ii = i + sh;
        ii = ii >= ns ? ii - ns : ii;
        const cx<TU> v = ldx(u, (long)ii * ns + j);
        cd x{(double)v.x, (double)v.y};
        if (ph) x = cmul(phs[i], x);
        return x;
    };
    for (int j = blockIdx.y * blockDim.x + threadIdx.x; j < ns; j += gridDim.y * blockDim.x) {
        cd a1{0.0, 0.0}, a2{0.0, 0.0};
        for (int i = 0; i < ns; i++) {
            const cd x = comp(u0, sh0, ph0, i, j);
            if (has1) cfmac(a1, x, comp(u1, sh1, ph1, i, j));
            if (has2) cfmac(a2, x, comp(u2, sh2, ph2, i, j));
        }
        st(Lk1, kp * ns + j, a1);
        st(Lk2, kp * ns + j, a2);
    }
}

// The same links for a tile of B1 x B2 points per workgroup: the tile's points and its
// neighbours below and to the right, (B1 + 1)(B2 + 1) - 1 vector sets, are read once per
// component row and every link of the tile is accumulated from registers, so a vector set
// is fetched (B1 + 1)(B2 + 1) - 1 times per B1 B2 points instead of 3 times per point.
// Grid (n_samples * nt1 * nt2, state chunks of 256), block 256, ns complex doubles of
// dynamic LDS; nt1 = ceil((g + 1) / B1), nt2 = ceil((g + 1) / B2). Each link is summed over
// the components in the same order as links_derived_kernel, so the results are identical.
template <typename TU, int B1, int B2>
__global__ __launch_bounds__(256) void links_tile_kernel(const TU *__restrict__ U, long strideU, int ns, int g,
                                                         int nt1, int nt2, double *__restrict__ Lk1,
                                                         double *__restrict__ Lk2) {
    extern __shared__ cd phs[];
    const int g1 = g + 1;
    const int tile = (int)(blockIdx.x % (unsigned)(nt1 * nt2));
    const long s = blockIdx.x / (unsigned)(nt1 * nt2);
    const int a0 = (tile / nt2) * B1, b0 = (tile % nt2) * B2;
    const double TWO_PI = 6.283185307179586232;
    if (a0 + B1 >= g) {                                      // tiles that touch theta_1 = 2 pi
        for (int i = threadIdx.x; i < ns; i += blockDim.x) {
            double sn, cs;
            sincos(TWO_PI * (double)i / (double)ns, &sn, &cs);
            phs[i] = cd{cs, sn};
        }
    }
    __syncthreads();
    constexpr int NS1 = B1 + 1, NS2 = B2 + 1;
    // slot (da, db): point (a0 + da, b0 + db) of the (g + 1)^2 mesh, mapped to its solved
    // set (theta = 2 pi rows and columns from theta = 0 with a component phase / shift);
    // the row part of its offset, validity and phase depend on da only, the column part and
    // the shift on db only
    const TU *base = U + 2 * (s * g * g) * strideU;
    long roff[NS1], coff[NS2];
    bool rok[NS1], rph[NS1], cok[NS2], csh[NS2];
    // rows and columns beyond the mesh keep the offset of the last one inside it (a0, b0 <=
    // g), so every load below reads a valid set and needs no branch; their values are unused
#pragma unroll
    for (int da = 0; da < NS1; da++) {
        const int a = a0 + da;
        rok[da] = a <= g;
        rph[da] = a == g;
        roff[da] = rok[da] ? 2 * (long)(a == g ? 0 : a) * g * strideU : roff[da > 0 ? da - 1 : 0];
    }
#pragma unroll
    for (int db = 0; db < NS2; db++) {
        const int b = b0 + db;
        cok[db] = b <= g;
        csh[db] = b == g;
        coff[db] = cok[db] ? 2 * (long)(b == g ? 0 : b) * strideU : coff[db > 0 ? db - 1 : 0];
    }
    auto ok = [&](int da, int db) { return rok[da] && cok[db] && !(da == B1 && db == B2); };
    for (int j = blockIdx.y * blockDim.x + threadIdx.x; j < ns; j += gridDim.y * blockDim.x) {
        cd a1[B1][B2], a2[B1][B2];
#pragma unroll
        for (int da = 0; da < B1; da++)
#pragma unroll
            for (int db = 0; db < B2; db++) {
                a1[da][db] = cd{0.0, 0.0};
                a2[da][db] = cd{0.0, 0.0};
            }
        for (int i = 0; i < ns; i++) {
            const long o0 = (long)i * ns + j;
            const long o1 = (long)(i + 1 >= ns ? i + 1 - ns : i + 1) * ns + j;
            // every load of the component row first (no branches between them), then the
            // links row by row: those to the row below use the previous row's values
            cx<TU> raw[NS1][NS2];
#pragma unroll
            for (int da = 0; da < NS1; da++)
#pragma unroll
                for (int db = 0; db < NS2; db++)
                    if (!(da == B1 && db == B2)) raw[da][db] = ldx(base + roff[da] + coff[db], csh[db] ? o1 : o0);
            cd xp[NS2];
#pragma unroll
            for (int da = 0; da < NS1; da++) {
                cd xr[NS2];
#pragma unroll
                for (int db = 0; db < NS2; db++) {
                    if (da == B1 && db == B2) {
                        xr[db] = cd{0.0, 0.0};
                        continue;
                    }
                    xr[db] = cd{(double)raw[da][db].x, (double)raw[da][db].y};
                    if (rph[da]) xr[db] = cmul(phs[i], xr[db]);
                }
                if (da > 0) {
#pragma unroll
                    for (int db = 0; db < B2; db++)
                        if (ok(da, db)) cfmac(a1[da - 1][db], xp[db], xr[db]);
                }
                if (da < B1) {
#pragma unroll
                    for (int db = 0; db < B2; db++)
                        if (ok(da, db + 1)) cfmac(a2[da][db], xr[db], xr[db + 1]);
                }
#pragma unroll
                for (int db = 0; db < NS2; db++) xp[db] = xr[db];
            }
        }
#pragma unroll
        for (int da = 0; da < B1; da++)
#pragma unroll
            for (int db = 0; db < B2; db++) {
                const int a = a0 + da, b = b0 + db;
                if (a > g || b > g) continue;
                const long kp = (s * g1 + a) * g1 + b;
                st(Lk1, kp * ns + j, a1[da][db]);
                st(Lk2, kp * ns + j, a2[da][db]);
            }
    }
}

// plaquettes + energies: grid (S), block 256 (states).
// mode 0 = rlm (+1/n_phi, chern += t), mode 1 = hlm (chern -= t), mode 2 = rlm with
// eigenvalues stored for the solved points only (theta < g; the theta = 2 pi points
// have the spectra of their theta = 0 counterparts), kp_s = (s*g + t1)*g + t2.
__global__ __launch_bounds__(256) void reduce_kernel(const double *__restrict__ Lk1, const double *__restrict__ Lk2,
                                                     const double *__restrict__ Wv, int ns, int n_mesh, int L2n,
                                                     int mode, double n_phi_d, double *__restrict__ energy,
                                                     double *__restrict__ chern, double *__restrict__ margin) {
    const int s = blockIdx.x;
    const int g = n_mesh, g1 = n_mesh + 1;
    const long kbase = (long)s * g1 * L2n;
    const double TWO_PI = 6.283185307179586232;   // (M_PI*2.0)
    const double PI = 3.141592653589793116;
    for (int j = threadIdx.x; j < ns; j += blockDim.x) {
        // energies (reference order: line by line, theta_2 ascending; the loads of a line
        // are issued before its additions, which keep that order)
        double e = 0.0;
        for (int t1 = 0; t1 <= g; t1++) {
            const long lb = kbase + (long)t1 * L2n;
            if (mode == 0) {
#pragma unroll 8
                for (int t2 = 0; t2 <= g; t2++) e += Wv[(lb + t2) * ns + j];
            } else if (mode == 2) {
                const long sb = ((long)s * g + (t1 % g)) * g;
#pragma unroll 8
                for (int t2 = 0; t2 <= g; t2++) e += Wv[(sb + t2 % g) * ns + j];
            } else {
#pragma unroll 8
                for (int t2 = 0; t2 < g; t2++) e += Wv[(lb + t2) * ns + j];
                e += Wv[lb * ns + j];
            }
        }
        e /= (double)(g1 * g1);
        // plaquettes (a = theta_1 - 1, b = m) in the reference's order, four at a time: the
        // sixteen links of four plaquettes load together, then their phases are taken in
        // order (one load latency per four plaquettes)
        auto links = [&](int a, int b, cd &l1, cd &l2, cd &l3, cd &l4) {
            const long k00 = kbase + (long)a * L2n + (b % L2n);
            const long k10 = k00 + L2n;
            const long k01 = kbase + (long)a * L2n + ((b + 1) % L2n);
            l1 = ld(Lk1, k00 * ns + j);     // <00|10>
            l2 = ld(Lk2, k10 * ns + j);     // <10|11>
            l3 = ld(Lk1, k01 * ns + j);     // conj: <11|01>
            l4 = ld(Lk2, k00 * ns + j);     // conj: <01|00>
        };
        double c = 0.0, mg = 4.0;
        for (int a = 0; a < g; a++) {
            for (int b0 = 0; b0 < g; b0 += 4) {
                cd L[4][4];
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (b0 + u < g) links(a, b0 + u, L[u][0], L[u][1], L[u][2], L[u][3]);
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (b0 + u >= g) break;
                    cd prod{1.0, 0.0};
                    prod = cmul(prod, L[u][0]);
                    prod = cmul(prod, L[u][1]);
                    prod = cmul(prod, cconj(L[u][2]));
                    prod = cmul(prod, cconj(L[u][3]));
                    const double ph = atan2(prod.y, prod.x);
                    const double tth = -ph / TWO_PI;
                    if (mode != 1) c += tth; else c += -tth;
                    mg = fmin(mg, PI - fabs(ph));
                }
            }
        }
        if (mode != 1) c += 1.0 / n_phi_d;
        energy[(long)s * ns + j] = e;
        chern[(long)s * ns + j] = c;
        if (margin) margin[(long)s * ns + j] = mg;
    }
}

int launch_links(const double *U, long strideU, int dim_vec, int ns, int lines, int L2n, int nlink2, long kp0,
                 long nkp, double *Lk1, double *Lk2, hipStream_t stream) {
    if (nkp <= 0) return 0;
    hipLaunchKernelGGL(links_kernel<double>, dim3((unsigned)nkp, (ns + 255) / 256), dim3(256), 0, stream, U, strideU,
                       dim_vec, ns, lines, L2n, nlink2, kp0, Lk1, Lk2);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_links_f32(const float *U, long strideU, int dim_vec, int ns, int lines, int L2n, int nlink2, long kp0,
                     long nkp, double *Lk1, double *Lk2, hipStream_t stream) {
    if (nkp <= 0) return 0;
    hipLaunchKernelGGL(links_kernel<float>, dim3((unsigned)nkp, (ns + 255) / 256), dim3(256), 0, stream, U, strideU,
                       dim_vec, ns, lines, L2n, nlink2, kp0, Lk1, Lk2);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_links_derived(const void *U, bool fp32, long strideU, int ns, int g, int n_samples, double *Lk1,
                         double *Lk2, hipStream_t stream) {
    const long nkp = (long)n_samples * (g + 1) * (g + 1);
    if (nkp <= 0) return 0;
    if (ns > 4096 || g < 1) return -22;   // phase table: ns complex doubles of LDS
    const size_t lds = sizeof(cd) * ns;
    // tiles of 3 x 3 points; FP64 vectors take 4 x 4 where that still gives at least four
    // workgroups per CU (n_phi = 256, g = 32, 32 samples: FP32 7.2 -> 4.3 ms, FP64 11.6 ->
    // 9.4 / 8.0 ms for 3 x 3 / 4 x 4; n_phi = 3000, g = 4: 6.3 -> 3.6 ms for 3 x 3, 5.0 ms
    // for 4 x 4, which leaves 192 workgroups). RLM_LINKS_TILE=0 keeps one point per
    // workgroup (links_derived_kernel), for comparison.
    const char *te = getenv("RLM_LINKS_TILE");
    if (!(te && atoi(te) == 0)) {
        const unsigned chunks = (unsigned)((ns + 255) / 256);
        int B = 3;
        if (!fp32) {
            int dev = 0, cus = 0;
            RLM_HIP_CHECK(hipGetDevice(&dev));
            RLM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            const long nt4 = (g + 1 + 3) / 4;
            if ((long)n_samples * nt4 * nt4 * chunks >= 4L * cus) B = 4;
        }
        const int nt = (g + 1 + B - 1) / B;
        const dim3 tgrid((unsigned)((long)n_samples * nt * nt), chunks);
        if (fp32)
            hipLaunchKernelGGL((links_tile_kernel<float, 3, 3>), tgrid, dim3(256), lds, stream,
                               static_cast<const float *>(U), strideU, ns, g, nt, nt, Lk1, Lk2);
        else if (B == 4)
            hipLaunchKernelGGL((links_tile_kernel<double, 4, 4>), tgrid, dim3(256), lds, stream,
                               static_cast<const double *>(U), strideU, ns, g, nt, nt, Lk1, Lk2);
        else
            hipLaunchKernelGGL((links_tile_kernel<double, 3, 3>), tgrid, dim3(256), lds, stream,
                               static_cast<const double *>(U), strideU, ns, g, nt, nt, Lk1, Lk2);
        RLM_HIP_CHECK(hipGetLastError());
        return 0;
    }
    const dim3 grid((unsigned)(8 * ((nkp + 7) / 8)), (ns + 255) / 256);
    if (fp32)
        hipLaunchKernelGGL(links_derived_kernel<float>, grid, dim3(256), lds, stream, static_cast<const float *>(U),
                           strideU, ns, g, nkp, Lk1, Lk2);
    else
        hipLaunchKernelGGL(links_derived_kernel<double>, grid, dim3(256), lds, stream, static_cast<const double *>(U),
                           strideU, ns, g, nkp, Lk1, Lk2);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_reduce(const double *Lk1, const double *Lk2, const double *Wv, int ns, int n_mesh, int L2n, int mode,
                  double n_phi_d, int n_samples, double *energy, double *chern, double *margin, hipStream_t stream) {
    hipLaunchKernelGGL(reduce_kernel, dim3(n_samples), dim3(256), 0, stream, Lk1, Lk2, Wv, ns, n_mesh, L2n, mode,
                       n_phi_d, energy, chern, margin);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- ring-buffer entry points
// One theta_1 line of the reference's 2-slot ring (diag_wrappers/chern.cpp:3-35): vertex
// (slot, theta_2) holds n_phi state vectors of dim_vec components, state-major
// (ring[((slot*(g+1) + theta_2)*n_phi + n)*dim_vec + i], the column-major eigenvector
// matrices the reference's solvers leave). One wave per (plaquette m, state n); the
// four overlaps follow the reference's path (slot 1 - theta_1 % 2, m) -> ... and its
// product order. io_order: the order of cal_Chern_wfs_IO (chern.cpp:111-118), factors
// <p_{l+1}|p_l> and +arg / 2 pi, instead of <p_l|p_{l+1}> and -arg / 2 pi.
// grid ceil(n_mesh * n_phi / 4), block 256; no barriers (waves past the end return).
__global__ __launch_bounds__(256) void ring_plaquette_kernel(const double *__restrict__ ring, int n_phi,
                                                             int dim_vec, int n_mesh, int theta_1, int io_order,
                                                             double *__restrict__ chern_theta) {
    const int lane = threadIdx.x & (kWave - 1);
    const long pair = (long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (pair >= (long)n_mesh * n_phi) return;
    const int m = (int)(pair / n_phi), n = (int)(pair % n_phi);
    const long dim_wfs = (long)n_phi * dim_vec;
    int dki[4], dkj[4];
    dki[2] = (theta_1 % 2 == 0) ? 1 : -1;
    dki[0] = -dki[2];
    dki[1] = dki[3] = 0;
    dkj[0] = dkj[2] = 0;
    dkj[3] = -1;
    dkj[1] = 1;
    int lki = 1 - theta_1 % 2, lkj = m;
    cd prod{1.0, 0.0};
    for (int l = 0; l < 4; l++) {
        const int rki = lki + dki[l], rkj = lkj + dkj[l];
        const double *ul = ring + 2 * (((long)lki * (n_mesh + 1) + lkj) * dim_wfs + (long)dim_vec * n);
        const double *ur = ring + 2 * (((long)rki * (n_mesh + 1) + rkj) * dim_wfs + (long)dim_vec * n);
        cd a{0.0, 0.0};
        for (int i = lane; i < dim_vec; i += kWave) {
            if (io_order) cfmac(a, ld(ur, i), ld(ul, i));   // <p_{l+1}|p_l>
            else cfmac(a, ld(ul, i), ld(ur, i));            // <p_l|p_{l+1}>
        }
        a.x = wave_sum(a.x);
        a.y = wave_sum(a.y);
        prod = cmul(prod, a);
        lki = rki;
        lkj = rkj;
    }
    if (lane == 0) {
        const double TWO_PI = 6.283185307179586232;
        const double ph = atan2(prod.y, prod.x);
        chern_theta[m + (long)n * n_mesh] = io_order ? ph / TWO_PI : -ph / TWO_PI;
    }
}

// Eigenvectors of line t1 of sample s (theta_2 = 0..g) in the reference's ring layout
// (state-major, out[(t2*ns + j)*ns + i] = component i of state j), as complex double or
// float. U holds the solved points, component-major: pl = g + 1 (every point solved,
// kp = (s*(g+1) + t1)*(g+1) + t2) or pl = g (theta < g solved, kp = (s*g + t1)*g + t2,
// the theta = 2 pi points derived as links_derived_kernel derives them).
// grid (g + 1, ceil(ns*ns / 256)), block 256.
template <typename TU, typename TO>
__global__ __launch_bounds__(256) void gather_line_kernel(const TU *__restrict__ U, long strideU, int ns, int g,
                                                          int pl, long s, int t1, TO *__restrict__ out) {
    const int t2 = blockIdx.x;
    const long e = (long)blockIdx.y * blockDim.x + threadIdx.x;
    if (e >= (long)ns * ns) return;
    const int j = (int)(e / ns), i = (int)(e % ns);
    int a = t1, b = t2, sh = 0;
    bool ph = false;
    if (pl == g) {
        sh = (b == g) ? 1 : 0;
        ph = (a == g);
        a = (a == g) ? 0 : a;
        b = (b == g) ? 0 : b;
    }
    const TU *base = U + 2 * (((s * pl + a) * pl + b) * strideU);
    int ii = i + sh;
    ii = ii >= ns ? ii - ns : ii;
    const cx<TU> v = ldx(base, (long)ii * ns + j);
    cd x{(double)v.x, (double)v.y};
    if (ph) {
        const double TWO_PI = 6.283185307179586232;
        double sn, cs;
        sincos(TWO_PI * (double)i / (double)ns, &sn, &cs);
        x = cmul(cd{cs, sn}, x);
    }
    TO *o = out + 2 * (((long)t2 * ns + j) * ns + i);
    o[0] = (TO)x.x;
    o[1] = (TO)x.y;
}

int launch_ring_plaquettes(const double *ring, int n_phi, int dim_vec, int n_mesh, int theta_1, int io_order,
                           double *chern_theta, hipStream_t stream) {
    if (n_phi < 1 || dim_vec < 1 || n_mesh < 1 || theta_1 < 1) return -22;
    const long pairs = (long)n_mesh * n_phi;
    hipLaunchKernelGGL(ring_plaquette_kernel, dim3((unsigned)((pairs + 3) / 4)), dim3(256), 0, stream, ring, n_phi,
                       dim_vec, n_mesh, theta_1, io_order, chern_theta);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_gather_line(const void *U, bool fp32_in, long strideU, int ns, int g, int pl, long s, int t1, void *out,
                       bool fp32_out, hipStream_t stream) {
    if (ns < 1 || g < 1 || (pl != g && pl != g + 1) || t1 < 0 || t1 > g) return -22;
    const dim3 grid((unsigned)(g + 1), (unsigned)(((long)ns * ns + 255) / 256));
    if (fp32_in && fp32_out)
        hipLaunchKernelGGL((gather_line_kernel<float, float>), grid, dim3(256), 0, stream, static_cast<const float *>(U),
                           strideU, ns, g, pl, s, t1, static_cast<float *>(out));
    else if (fp32_in)
        hipLaunchKernelGGL((gather_line_kernel<float, double>), grid, dim3(256), 0, stream,
                           static_cast<const float *>(U), strideU, ns, g, pl, s, t1, static_cast<double *>(out));
    else if (fp32_out)
        hipLaunchKernelGGL((gather_line_kernel<double, float>), grid, dim3(256), 0, stream,
                           static_cast<const double *>(U), strideU, ns, g, pl, s, t1, static_cast<float *>(out));
    else
        hipLaunchKernelGGL((gather_line_kernel<double, double>), grid, dim3(256), 0, stream,
                           static_cast<const double *>(U), strideU, ns, g, pl, s, t1, static_cast<double *>(out));
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace rlm
