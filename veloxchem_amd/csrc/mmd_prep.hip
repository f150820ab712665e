// Hermite-to-spherical matrices of the matrix-core MD Fock kernel, built once
// per pair set: U[prim][h][s] = k sum_{cart a, b} S_a S_b E^x_t E^y_u E^z_v for
// Hermite index h = (t, u, v) of degree <= la + lb and spherical index
// s = (m_a, m_b) of the pair; zero in the padding rows and columns.
#include "md_kernels.h"

namespace {

__global__ void __launch_bounds__(VLX_WAVE)
    u_build_kernel(const PairInfo* __restrict__ pairs, const PrimPair* __restrict__ prims, int npairs, int la, int lb,
                   double* __restrict__ U)
{
    __shared__ double e[3 * md::E3];
    const int nh = md::nherm(la + lb), nsb = 2 * lb + 1, ns = (2 * la + 1) * nsb;
    const int rows = vlx_u_rows(la, lb), cols = vlx_u_cols(la, lb);
    for (int pi = blockIdx.x; pi < npairs; pi += gridDim.x) {
        const PairInfo pp = pairs[pi];
        for (int q = 0; q < pp.nprim; q++) {
            const PrimPair pr = prims[pp.prim_start + q];
            const double pa[3] = {pr.pax, pr.pay, pr.paz};
            const double pb[3] = {pr.pax + pp.abx, pr.pay + pp.aby, pr.paz + pp.abz};
            __syncthreads();
            md::hermite_1d(la, lb, pr.zeta, pa, pb, e);
            __syncthreads();
            double* u = U + (size_t)(pp.prim_start + q) * rows * cols;
            for (int idx = threadIdx.x; idx < rows * cols; idx += blockDim.x) {
                const int h = idx / cols, sidx = idx % cols;
                double acc = 0.0;
                if (h < nh && sidx < ns) {
                    int t, uu, v;
                    md::hdecode(h, t, uu, v);
                    int cia[3], cib[3];
                    double cfa[3], cfb[3];
                    const int nta = md::sph_terms(la, sidx / nsb, cia, cfa);
                    const int ntb = md::sph_terms(lb, sidx % nsb, cib, cfb);
                    for (int ia = 0; ia < nta; ia++) {
                        int ax, ay, az;
                        md::cart_exp(la, cia[ia], ax, ay, az);
                        for (int ib = 0; ib < ntb; ib++) {
                            int bx, by, bz;
                            md::cart_exp(lb, cib[ib], bx, by, bz);
                            if (t > ax + bx || uu > ay + by || v > az + bz) continue;
                            acc += cfa[ia] * cfb[ib] * e[(ax * md::D1 + bx) * md::DT + t] *
                                   e[md::E3 + (ay * md::D1 + by) * md::DT + uu] *
                                   e[2 * md::E3 + (az * md::D1 + bz) * md::DT + v];
                        }
                    }
                    acc *= pr.k;
                }
                u[idx] = acc;
            }
        }
    }
}

}  // namespace

void vlx_launch_u_build(const PairInfo* pairs, const PrimPair* prims, int npairs, int la, int lb, double* u,
                        hipStream_t s)
{
    if (npairs <= 0) return;
    const int g = std::min(npairs, 16384);
    hipLaunchKernelGGL(u_build_kernel, dim3(g), dim3(VLX_WAVE), 0, s, pairs, prims, npairs, la, lb, u);
}
