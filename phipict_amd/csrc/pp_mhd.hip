// Cell-wise cross products of the inductionless MHD step, one launch each instead of a chain
// of tensor operations: u x e_b for the potential's right-hand side and the current's input
// (src/phipict/simulation/mhd.py:276-355), and the Lorentz force N (J x e_b) added to a
// block's velocity source (mhd.py:543-594).
#include "pp_common.hpp"

#include <algorithm>

namespace pp {

// (a x b)_c of 3-vectors, component by component as the reference's cross product
template <typename T>
__device__ __forceinline__ T cross_c(int c, T a0, T a1, T a2, T b0, T b1, T b2) {
    return c == 0 ? a1 * b2 - a2 * b1 : (c == 1 ? a2 * b0 - a0 * b2 : a0 * b1 - a1 * b0);
}

// out[b][c][i] = (v x e_b)_c for c < 3; v [B][vdims][n] lifted to 3D with zero components
template <typename T>
__global__ void __launch_bounds__(256) k_cross_eb(const T *__restrict__ v, int vdims, int n, const T *__restrict__ eb,
                                                  int ebPerBatch, T *__restrict__ out) {
    const int b = blockIdx.y;
    const T *e = eb + (ebPerBatch ? 3 * b : 0);
    const T e0 = e[0], e1 = e[1], e2 = e[2];
    const T *vb = v + (size_t)b * vdims * n;
    T *ob = out + (size_t)b * 3 * n;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const T a0 = vb[i], a1 = vdims > 1 ? vb[(size_t)n + i] : T(0), a2 = vdims > 2 ? vb[(size_t)2 * n + i] : T(0);
        ob[i] = cross_c<T>(0, a0, a1, a2, e0, e1, e2);
        ob[(size_t)n + i] = cross_c<T>(1, a0, a1, a2, e0, e1, e2);
        ob[(size_t)2 * n + i] = cross_c<T>(2, a0, a1, a2, e0, e1, e2);
    }
}

// out[b][c][f] = N_b (J(g0 + f) x e_b)_c + S, c < dims, f < cells; J [B][3][total];
// S: none, static [sb][dims] or per cell [sb][dims][cells] (sb = 1 or B)
template <typename T>
__global__ void __launch_bounds__(256) k_lorentz_source(const T *__restrict__ J, int total, int dims,
                                                        const T *__restrict__ eb, int ebPerBatch,
                                                        const T *__restrict__ N, int nPerBatch, int g0, int cells,
                                                        const T *__restrict__ src, int srcKind, int srcBatched,
                                                        T *__restrict__ out) {
    const int b = blockIdx.y;
    const T *e = eb + (ebPerBatch ? 3 * b : 0);
    const T e0 = e[0], e1 = e[1], e2 = e[2];
    const T nb = N[nPerBatch ? b : 0];
    const T *Jb = J + (size_t)b * 3 * total + g0;
    const int sb = srcBatched ? b : 0;
    T *ob = out + (size_t)b * dims * cells;
    for (int f = blockIdx.x * 256 + threadIdx.x; f < cells; f += gridDim.x * 256) {
        const T a0 = Jb[f], a1 = Jb[(size_t)total + f], a2 = Jb[(size_t)2 * total + f];
        for (int c = 0; c < dims; ++c) {
            T F = nb * cross_c<T>(c, a0, a1, a2, e0, e1, e2);
            if (srcKind == 1) F = F + src[(size_t)sb * dims + c];
            else if (srcKind == 2) F = F + src[((size_t)sb * dims + c) * cells + f];
            ob[(size_t)c * cells + f] = F;
        }
    }
}

static dim3 cell_grid(int n, int batch) { return dim3(std::max(1, std::min((n + 255) / 256, 4096)), batch); }
} // namespace pp

using namespace pp;

extern "C" {
int pp_cross_eb(const void *v, int32_t vdims, int32_t n, int32_t batch, const void *e_b, int32_t eb_per_batch, void *out,
                int32_t dtype, void *stream) {
    PP_CHECK_ARG(v && e_b && out, "null argument");
    PP_CHECK_ARG(vdims >= 1 && vdims <= 3 && n >= 0 && batch >= 1, "bad sizes");
    if (n == 0) return PP_OK;
    PP_DISPATCH(dtype, hipLaunchKernelGGL((k_cross_eb<T>), cell_grid(n, batch), dim3(256), 0, as_stream(stream),
                                          static_cast<const T *>(v), (int)vdims, (int)n, static_cast<const T *>(e_b),
                                          (int)eb_per_batch, static_cast<T *>(out)));
    PP_LAUNCH_CHECK();
    return PP_OK;
}

int pp_lorentz_source(const void *J, int32_t total, int32_t batch, int32_t dims, const void *e_b, int32_t eb_per_batch,
                      const void *stuart, int32_t stuart_per_batch, int32_t block_offset, int32_t block_cells,
                      const void *src, int32_t src_kind, int32_t src_batched, void *out, int32_t dtype, void *stream) {
    PP_CHECK_ARG(J && e_b && stuart && out, "null argument");
    PP_CHECK_ARG(dims >= 1 && dims <= 3 && batch >= 1 && total >= 0 && block_cells >= 0, "bad sizes");
    PP_CHECK_ARG(block_offset >= 0 && (int64_t)block_offset + block_cells <= total, "block outside the domain");
    PP_CHECK_ARG(src_kind >= 0 && src_kind <= 2 && (src_kind == 0 || src), "bad source");
    if (block_cells == 0) return PP_OK;
    PP_DISPATCH(dtype, hipLaunchKernelGGL((k_lorentz_source<T>), cell_grid(block_cells, batch), dim3(256), 0,
                                          as_stream(stream), static_cast<const T *>(J), (int)total, (int)dims,
                                          static_cast<const T *>(e_b), (int)eb_per_batch, static_cast<const T *>(stuart),
                                          (int)stuart_per_batch, (int)block_offset, (int)block_cells,
                                          static_cast<const T *>(src), (int)src_kind, (int)src_batched,
                                          static_cast<T *>(out)));
    PP_LAUNCH_CHECK();
    return PP_OK;
}
}
