// Patch form of the matrix-free operator y = (cs S + ct T + i c0 S0) x.
//
// The per-cell form (matfree.hip) writes every element product to HBM and reads it back
// in a second kernel that sums each DOF's products.  Here the cells are ordered along a
// Morton curve of their centroids and cut into patches of CB consecutive cells (compact
// clusters of the mesh; CB = 128, 192 or 256, chosen by the autotuner), one workgroup
// per patch:
//
//  1. the patch's DOFs ("patch DOFs": the union of its cells' DOFs, ~1.4 per DOF of the
//     mesh) are gathered from x into LDS once;
//  2. each thread forms its cell's element product from the LDS copy (the same
//     instruction-folded element matrix as the per-cell form), the boundary faces of the
//     patch their ABC products;
//  3. every product is stored in LDS at its place in the patch's DOF-sorted order, and
//     each patch DOF's segment is summed in that fixed order.  A DOF that lies in this
//     patch only (about two thirds of them) is finished: its row of y is written (PEC row
//     mask and the Krylov epilogue applied here).  The partial sum of a DOF shared with
//     other patches goes to its slot in a buffer sorted by (DOF, patch);
//  4. k_mf_psum adds each shared DOF's partial sums in patch order, then the mask and the
//     epilogue.
//
// HBM traffic per application: 12 coefficients and nloc (patch-DOF index, product place)
// pairs per cell; per patch DOF its id, segment end and destination; the gathered x; y
// once; the shared DOFs' partial sums written and read once -- no per-product round
// trip.  Every sum has a fixed order, so results are deterministic (and independent of
// the launch shape of step 4).
#include <algorithm>
#include <cstring>
#include <vector>
#include <rocprim/rocprim.hpp>

#include "internal.hpp"
#include "kernels.hpp"
#include "ops.hpp"
#include "mf_element.hpp"

namespace sucem {
namespace {

constexpr int TB = 256;    // threads of the set-up kernels and of k_mf_psum
// patch DOFs per thread and batch in k_mf_patch: of the x gather (UB) and of the segment
// sums (UD).  Patches of 256 cells hold about 2,300 DOFs at order 2, so 10 per thread
// takes them in one batch; measured per FP32 application at 40^3, order 2: 8 / 10 / 12 /
// 14 / 16 -> 64.6 / 59.9 / 60.5 / 60.5 / 62.0 us.
constexpr int UB = 10;
constexpr int UD = 10;
constexpr int kCbs[3] = {128, 192, 256};   // cells per patch = threads of k_mf_patch

int cb_slot(int cb) { return cb == 128 ? 0 : (cb == 192 ? 1 : (cb == 256 ? 2 : -1)); }

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 30-bit Morton code of each cell's centroid on a 1024^3 grid over the mesh's box
__global__ void k_pt_morton(int64_t nc, const double* xyz, const int32_t* cells, double lx,
                            double ly, double lz, double sx, double sy, double sz, uint32_t* key,
                            uint32_t* val) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc;
       c += (int64_t)gridDim.x * blockDim.x) {
    double p[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < 4; ++v) {
      const int64_t i = cells[4 * c + v];
      p[0] += xyz[3 * i];
      p[1] += xyz[3 * i + 1];
      p[2] += xyz[3 * i + 2];
    }
    const double lo[3] = {lx, ly, lz}, sc[3] = {sx, sy, sz};
    uint32_t q[3];
    for (int d = 0; d < 3; ++d) {
      const double t = (0.25 * p[d] - lo[d]) * sc[d];
      q[d] = t <= 0.0 ? 0u : (t >= 1023.0 ? 1023u : (uint32_t)t);
    }
    key[c] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
    val[c] = (uint32_t)c;
  }
}

__global__ void k_pt_rank(int64_t nc, const int32_t* order, int32_t* rank) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nc;
       r += (int64_t)gridDim.x * blockDim.x)
    rank[order[r]] = (int32_t)r;
}

__global__ void k_pt_face_keys(int64_t nbf, const int32_t* bf_cell, const int32_t* rank,
                               uint32_t* key, uint32_t* val) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nbf;
       f += (int64_t)gridDim.x * blockDim.x) {
    key[f] = (uint32_t)rank[bf_cell[f]];
    val[f] = (uint32_t)f;
  }
}

// faces in patch order: (cell rank << 2) | local face
__global__ void k_pt_face_meta(int64_t nbf, const uint32_t* rank_sorted, const int32_t* face,
                               const int32_t* bf_local, int32_t* meta) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nbf;
       i += (int64_t)gridDim.x * blockDim.x)
    meta[i] = (int32_t)((rank_sorted[i] << 2) | (uint32_t)bf_local[face[i]]);
}

__global__ void k_pt_meta_rank(int64_t nbf, const int32_t* meta, uint32_t* rank) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nbf;
       i += (int64_t)gridDim.x * blockDim.x)
    rank[i] = (uint32_t)(meta[i] >> 2);
}

// ptr[q] = first i with grp(i) >= q (q = 0..ng), for group ids ascending in i; the
// group of entry i is key[i] >> shift, divided by div.
template <typename K>
__global__ void k_pt_ptr(int64_t n, const K* key, int shift, int64_t div, int64_t ng,
                         int32_t* ptr) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = (int64_t)(key[i] >> shift) / div;
    const int64_t lo = i == 0 ? 0 : (int64_t)(key[i - 1] >> shift) / div + 1;
    for (int64_t q = lo; q <= g; ++q) ptr[q] = (int32_t)i;
    if (i == n - 1)
      for (int64_t q = g + 1; q <= ng; ++q) ptr[q] = (int32_t)n;
  }
}

// Every product slot: cells e = r * nloc + l (r = rank), then the faces in patch order
// (nc * nloc + fp * nfd + a); key = (patch << 32) | DOF.
__global__ void k_pt_entries(int64_t nc, int nloc, int64_t nbf, int nfd, int cb, const int32_t* order,
                             const int32_t* cell_dofs, const int32_t* meta,
                             const int32_t* face_dofs, uint64_t* key, uint32_t* val) {
  const int64_t ncl = nc * nloc, n = ncl + nbf * nfd;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r;
    int l;
    if (e < ncl) {
      r = e / nloc;
      l = (int)(e % nloc);
    } else {
      const int64_t fp = (e - ncl) / nfd;
      const int a = (int)((e - ncl) % nfd);
      const int mt = meta[fp];
      r = mt >> 2;
      l = face_dofs[(mt & 3) * nfd + a];
    }
    const uint64_t dof = (uint64_t)cell_dofs[(int64_t)order[r] * nloc + l];
    key[e] = ((uint64_t)(r / cb) << 32) | dof;
    val[e] = (uint32_t)e;
  }
}

__global__ void k_pt_heads(int64_t n, const uint64_t* key, int32_t* flag) {
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < n;
       z += (int64_t)gridDim.x * blockDim.x)
    flag[z] = (z == 0 || key[z] != key[z - 1]) ? 1 : 0;
}

// uidx = inclusive scan of the head flags: patch DOF of sorted entry z is uidx[z] - 1
__global__ void k_pt_units(int64_t n, const uint64_t* key, const int32_t* uidx,
                           const int32_t* pstart, int32_t* pdof, uint16_t* pseg) {
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < n;
       z += (int64_t)gridDim.x * blockDim.x) {
    if (z == n - 1 || key[z + 1] != key[z]) {
      const int u = uidx[z] - 1;
      const int64_t p = (int64_t)(key[z] >> 32);
      pdof[u] = (int32_t)(key[z] & 0xffffffffu);
      pseg[u] = (uint16_t)(z - pstart[p] + 1);
    }
  }
}

__global__ void k_pt_uptr(int64_t np, int64_t n, const int32_t* pstart, const int32_t* uidx,
                          int32_t* uptr) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p <= np;
       p += (int64_t)gridDim.x * blockDim.x)
    uptr[p] = p < np ? uidx[pstart[p]] - 1 : uidx[n - 1];
}

__global__ void k_pt_places(int64_t n, int64_t nc, int nloc, const uint64_t* key,
                            const uint32_t* slot, const int32_t* uidx, const int32_t* pstart,
                            const int32_t* uptr, uint32_t* lp, uint16_t* fpos) {
  const int64_t ncl = nc * nloc;
  for (int64_t z = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; z < n;
       z += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = (int64_t)(key[z] >> 32);
    const uint32_t place = (uint32_t)(z - pstart[p]);
    const uint32_t lidx = (uint32_t)(uidx[z] - 1 - uptr[p]);
    const int64_t e = slot[z];
    if (e < ncl) lp[(e % nloc) * nc + e / nloc] = lidx | (place << 16);
    else fpos[e - ncl] = (uint16_t)place;
  }
}

__global__ void k_pt_count(int64_t nu, const int32_t* pdof, int32_t* cnt) {
  for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nu;
       u += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[pdof[u]], 1);
}

__global__ void k_pt_iota(int64_t n, uint32_t* v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (uint32_t)i;
}

// over the patch DOFs sorted by (DOF, patch): shared = the DOF lies in several patches;
// head = first entry of a shared DOF
__global__ void k_pt_shared(int64_t nu, const uint32_t* dk, const int32_t* cnt, int32_t* shared,
                            int32_t* head) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nu;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int sh = cnt[dk[i]] > 1 ? 1 : 0;
    shared[i] = sh;
    head[i] = (sh && (i == 0 || dk[i] != dk[i - 1])) ? 1 : 0;
  }
}

__global__ void k_pt_outs(int64_t nu, const uint32_t* dk, const uint32_t* du,
                          const int32_t* shared, const int32_t* ppos, const int32_t* head,
                          const int32_t* qi, int32_t* pout, int32_t* qdof, int32_t* qptr) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nu;
       i += (int64_t)gridDim.x * blockDim.x) {
    pout[du[i]] = shared[i] ? ppos[i] : -1;
    if (head[i]) {
      qdof[qi[i]] = (int32_t)dk[i];
      qptr[qi[i]] = ppos[i];
    }
  }
}

__global__ void k_set_i32_pt(int64_t n, int32_t* a, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void k_pt_qcode(int64_t npd, const int32_t* qptr, const int32_t* qdof, int32_t* qcode,
                           int32_t* too_many) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < npd;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int cnt = qptr[k + 1] - qptr[k];
    if (cnt > 15) atomicMax(too_many, 1);
    qcode[qdof[k]] = (qptr[k] << 4) | (cnt & 15);
  }
}

template <typename T>
__global__ void k_pt_permute(int64_t n, int k, const int32_t* from, const T* src, T* dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * k;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = i / n, r = i % n;
    dst[i] = src[j * n + from[r]];
  }
}

template <int NT>
__device__ __forceinline__ void block_partials(const RedSite& rs, int Gt, int at, double2 d1,
                                               double2 d2) {
  d1 = block_sum<NT>(d1);
  d2 = block_sum<NT>(d2);
  if (threadIdx.x == 0) {
    rs.part[at] = d1;
    rs.part[Gt + at] = d2;
  }
}

// EPI 1: w.y (bilinear), EPI 2: conj(w).y, EPI 4: w.y per part (paired systems)
template <int EPI, typename C>
__device__ __forceinline__ void epilogue(double2& d1, double2& d2, C v, C wv) {
  const double2 yd = to_d(v);
  double2 wd = to_d(wv);
  if constexpr (EPI == 4) {
    d1.x += wd.x * yd.x;
    d1.y += wd.y * yd.y;
    return;
  }
  if (EPI == 2) wd.y = -wd.y;
  d1.x += wd.x * yd.x - wd.y * yd.y;
  d1.y += wd.x * yd.y + wd.y * yd.x;
  d2.x += yd.x * yd.x + yd.y * yd.y;
}

// One workgroup per patch (XCD-aware order: consecutive patches share x entries and
// should share an L2); Gp = grid (a multiple of 8), Gt = all block partials of the
// operator (Gp + the grid of k_mf_psum).  EPI 3 (COCG, shared rows left as partial sums)
// reduces p.Ap; EPI 4 is EPI 3 for two paired real systems (per-part products).
template <typename R, int ORDER, int NFD, int EPI, int CB>
__global__ __launch_bounds__(CB) void k_mf_patch(
    int64_t np, int Gp, int Gt, int64_t nc, int64_t nbf, int lds_main, const R* __restrict__ coef,
    const uint32_t* __restrict__ lp, const int32_t* __restrict__ uptr,
    const int32_t* __restrict__ pdof, const uint16_t* __restrict__ pseg,
    const int32_t* __restrict__ pout, const int32_t* __restrict__ fptr,
    const int32_t* __restrict__ fmeta, const uint16_t* __restrict__ fpos,
    const R* __restrict__ hf, const int32_t* __restrict__ face_dofs,
    const double* __restrict__ ref_face, R cs, R ct, R c0, const uint8_t* __restrict__ mask,
    const typename C2<R>::T* __restrict__ x, typename C2<R>::T* __restrict__ y,
    typename C2<R>::T* __restrict__ part, const typename C2<R>::T* __restrict__ w,
    RedSite rs) {
  using C = typename C2<R>::T;
  constexpr bool E3 = EPI == 3 || EPI == 4;   // COCG form (EPI 4: paired systems)
  constexpr bool PR = EPI == 4;
  constexpr int NLOC = MfTab<ORDER>::NLOC;
  constexpr int KK = NFD * NFD;
  extern __shared__ double2 smem_[];
  C* sm = reinterpret_cast<C*>(smem_);
  const int lb = xcd_block(Gp);
  const int64_t p = lb;
  double2 d1 = make_double2(0.0, 0.0), d2 = make_double2(0.0, 0.0);
  if (p < np) {   // uniform over the workgroup
    const int u0 = uptr[p], nu = uptr[p + 1] - u0;
    const int64_t r = p * CB + threadIdx.x;
    const bool has = r < nc;
    // FP32: the cell's coefficients and LDS places do not depend on x, so their loads go
    // out with the x gather below and the two round trips overlap (in FP64 the held
    // registers would cost a wave per SIMD)
    constexpr bool PF = sizeof(R) == 4;
    R pa[12];
    uint32_t pli[NLOC];
    if (PF && has) {
#pragma unroll
      for (int k = 0; k < 12; ++k) pa[k] = coef[k * nc + r];
#pragma unroll
      for (int i = 0; i < NLOC; ++i) pli[i] = lp[(int64_t)i * nc + r];
    }
    // batches of UB entries per thread: all id loads, then all x loads, in flight together
    for (int j0 = 0; j0 < nu; j0 += UB * CB) {
      int dv[UB];
#pragma unroll
      for (int q = 0; q < UB; ++q) {
        const int j = j0 + q * CB + threadIdx.x;
        dv[q] = j < nu ? pdof[u0 + j] : -1;
      }
      C xv[UB];
#pragma unroll
      for (int q = 0; q < UB; ++q) xv[q] = dv[q] >= 0 ? x[dv[q]] : mk(R(0), R(0));
#pragma unroll
      for (int q = 0; q < UB; ++q) {
        const int j = j0 + q * CB + threadIdx.x;
        if (j < nu) sm[j] = xv[q];
      }
    }
    __syncthreads();
    C acc[NLOC];
    if (has) {
      R a[12];
      C xl[NLOC];
      if constexpr (PF) {
#pragma unroll
        for (int k = 0; k < 12; ++k) a[k] = pa[k] * (k < 6 ? ct : cs) * R(1.0 / MfTab<ORDER>::den);
#pragma unroll
        for (int i = 0; i < NLOC; ++i) xl[i] = sm[pli[i] & 0xffffu];
      } else {
#pragma unroll
        for (int k = 0; k < 12; ++k)
          a[k] = coef[k * nc + r] * (k < 6 ? ct : cs) * R(1.0 / MfTab<ORDER>::den);
#pragma unroll
        for (int i = 0; i < NLOC; ++i) xl[i] = sm[lp[(int64_t)i * nc + r] & 0xffffu];
      }
#pragma unroll
      for (int i = 0; i < NLOC; ++i) acc[i] = mk(R(0), R(0));
      element_product<R, ORDER>(a, xl, acc, std::make_integer_sequence<int, MfTab<ORDER>::NP>{});
      if constexpr (E3 && sizeof(R) == 8) {
        // p.Ap as the sum of the element quadratic forms x_c^T K_c x_c (w = x in COCG;
        // x is zero on masked rows there, so the masked rows add nothing either way).
        // FP64 only: in FP32 the extra registers cost a wave per SIMD, and the sums of
        // the finished rows and partial sums below are cheaper.
#pragma unroll
        for (int i = 0; i < NLOC; ++i) {
          const double2 xd = to_d(xl[i]), yd = to_d(acc[i]);
          if constexpr (PR) {
            d1.x += xd.x * yd.x;
            d1.y += xd.y * yd.y;
          } else {
            d1.x += xd.x * yd.x - xd.y * yd.y;
            d1.y += xd.x * yd.y + xd.y * yd.x;
          }
        }
      }
    }
    // the patch's boundary faces, one (face, row) pair per thread and pass, staged in LDS
    // behind the x copy (stage[id]) until the x copy may be overwritten
    const int f0 = fptr[p], nfr = (fptr[p + 1] - f0) * NFD;
    C* stage = sm + lds_main;
    for (int id = threadIdx.x; id < nfr; id += CB) {
      {
        const int64_t fp = f0 + id / NFD;
        const int a = id % NFD;
        const int mt = fmeta[fp];
        const int64_t rk = mt >> 2;
        const int lf = mt & 3;
        R hk[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) hk[k] = hf[k * nbf + fp];
        const double* rf = ref_face + lf * 6 * KK + a * NFD;
        C v = mk(R(0), R(0));
#pragma unroll
        for (int b = 0; b < NFD; ++b) {
          R m = R(0);
#pragma unroll
          for (int k = 0; k < 6; ++k) m += hk[k] * (R)rf[k * KK + b];
          const uint32_t li = lp[(int64_t)face_dofs[lf * NFD + b] * nc + rk] & 0xffffu;
          const C xb = sm[li];
          v.x += m * xb.x;
          v.y += m * xb.y;
        }
        stage[id] = mk(-c0 * v.y, c0 * v.x);
        if constexpr (E3 && sizeof(R) == 8) {
          const C xa = sm[lp[(int64_t)face_dofs[lf * NFD + a] * nc + rk] & 0xffffu];
          const double2 xd = to_d(xa), yd = to_d(stage[id]);
          if constexpr (PR) {
            d1.x += xd.x * yd.x;
            d1.y += xd.y * yd.y;
          } else {
            d1.x += xd.x * yd.x - xd.y * yd.y;
            d1.y += xd.x * yd.y + xd.y * yd.x;
          }
        }
      }
    }
    __syncthreads();   // every read of the x copy is done: the products overwrite it
    if (has) {
#pragma unroll
      for (int i = 0; i < NLOC; ++i) sm[lp[(int64_t)i * nc + r] >> 16] = acc[i];
    }
    int lo[UD], hi[UD], out[UD], dv[UD];
    auto load_ids = [&](int j0) {
#pragma unroll
      for (int q = 0; q < UD; ++q) {
        const int j = j0 + q * CB + threadIdx.x;
        const bool in = j < nu;
        lo[q] = (in && j) ? (int)pseg[u0 + j - 1] : 0;
        hi[q] = in ? (int)pseg[u0 + j] : 0;
        out[q] = in ? pout[u0 + j] : 0;
        dv[q] = in ? pdof[u0 + j] : 0;
      }
    };
    if constexpr (PF) load_ids(0);   // before the barrier: the loads overlap the wait
    for (int id = threadIdx.x; id < nfr; id += CB) sm[fpos[(int64_t)f0 * NFD + id]] = stage[id];
    __syncthreads();
    for (int j0 = 0; j0 < nu; j0 += UD * CB) {
      if (!PF || j0) load_ids(j0);
#pragma unroll
      for (int q = 0; q < UD; ++q) {
        C v = mk(R(0), R(0));
        for (int t = lo[q]; t < hi[q]; ++t) {
          const C a = sm[t];
          v.x += a.x;
          v.y += a.y;
        }
        if (hi[q] > 0) {   // an entry of this patch (segments are never empty)
          if (out[q] >= 0) {
            part[out[q]] = v;
            // FP32: x.y over the partial sums (linear in y; x is zero on masked rows)
            if constexpr (E3 && sizeof(R) == 4) epilogue<PR ? 4 : 1>(d1, d2, v, w[dv[q]]);
          } else {
            if (mask && mask[dv[q]]) v = mk(R(0), R(0));
            y[dv[q]] = v;
            if constexpr (EPI == 1 || EPI == 2) epilogue<EPI>(d1, d2, v, w[dv[q]]);
            else if constexpr (E3 && sizeof(R) == 4) epilogue<PR ? 4 : 1>(d1, d2, v, w[dv[q]]);
          }
        }
      }
    }
  }
  if (EPI) block_partials<CB>(rs, Gt, lb, d1, d2);
}

// Shared DOFs: y_d = sum of the partial sums of its patches, in patch order.
template <typename R, int EPI>
__global__ __launch_bounds__(TB) void k_mf_psum(int npd, int Gg, int Gt, int at0,
                                                const int32_t* __restrict__ qptr,
                                                const int32_t* __restrict__ qdof,
                                                const typename C2<R>::T* __restrict__ part,
                                                const uint8_t* __restrict__ mask,
                                                typename C2<R>::T* __restrict__ y,
                                                const typename C2<R>::T* __restrict__ w,
                                                RedSite rs) {
  using C = typename C2<R>::T;
  const int lb = xcd_block(Gg);
  int chunk = (npd + Gg - 1) / Gg;
  chunk = (chunk + TB - 1) / TB * TB;
  const int start = lb * chunk;
  const int end = min(npd, start + chunk);
  double2 d1 = make_double2(0.0, 0.0), d2 = make_double2(0.0, 0.0);
  for (int i = start + threadIdx.x; i < end; i += TB) {
    C v = mk(R(0), R(0));
    const int q1 = qptr[i + 1];
    for (int q = qptr[i]; q < q1; ++q) {
      const C t = part[q];
      v.x += t.x;
      v.y += t.y;
    }
    const int d = qdof[i];
    if (mask && mask[d]) v = mk(R(0), R(0));
    y[d] = v;
    if (EPI) epilogue<EPI>(d1, d2, v, w[d]);
  }
  if (EPI) block_partials<TB>(rs, Gt, at0 + lb, d1, d2);
}

template <typename T>
void scan(Context* c, const int32_t* in, int32_t* out, int64_t n, bool inclusive) {
  size_t bytes = 0;
  if (inclusive) {
    SUCEM_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, (size_t)n, rocprim::plus<int32_t>(),
                                      c->stream));
    DBuf<char> tmp(bytes);
    SUCEM_HIP(rocprim::inclusive_scan(tmp.p, bytes, in, out, (size_t)n, rocprim::plus<int32_t>(),
                                      c->stream));
    c->sync();
  } else {
    SUCEM_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n,
                                      rocprim::plus<int32_t>(), c->stream));
    DBuf<char> tmp(bytes);
    SUCEM_HIP(rocprim::exclusive_scan(tmp.p, bytes, in, out, 0, (size_t)n,
                                      rocprim::plus<int32_t>(), c->stream));
    c->sync();
  }
}

int32_t read_i32(Context* c, const int32_t* d) {
  int32_t h = 0;
  SUCEM_HIP(hipMemcpyAsync(&h, d, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  return h;
}

void pt_validate(Context* c, const Space* s, const PatchMaps& m, const std::vector<int32_t>& ps,
                 const std::vector<int32_t>& up, const std::vector<int32_t>& fp);

// Morton order of the cells and patch order of the boundary faces (shared by every patch
// size); s->pt_state = -1 when the space does not suit the patch kernel.
void pt_build_order(Context* c, Space* s) {
  const Mesh* m = s->mesh;
  const int64_t nc = m->nc, nbf = s->nbf;
  if (!(s->nloc == 20 && s->nfd == 8) && !(s->nloc == 6 && s->nfd == 3)) {
    s->pt_state = -1;
    return;
  }
  const int64_t ne = nc * s->nloc + nbf * s->nfd;
  SUCEM_CHECK(ne < (int64_t(1) << 31) && s->ndofs < (int64_t(1) << 31) && nc < (int64_t(1) << 29),
              SUCEM_ERR_UNSUPPORTED, "mesh too large for the patch operator");
  std::vector<double> xyz((size_t)m->nv * 3);
  d2h(c, xyz.data(), m->xyz.p, xyz.size());
  c->sync();
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = 0; i < m->nv; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], xyz[3 * i + d]);
      hi[d] = std::max(hi[d], xyz[3 * i + d]);
    }
  double sc[3];
  for (int d = 0; d < 3; ++d) sc[d] = hi[d] > lo[d] ? 1023.999 / (hi[d] - lo[d]) : 0.0;
  DBuf<int32_t> rank(nc);
  {
    DBuf<uint32_t> key(nc), key_s(nc), val(nc);
    s->pt_cell.alloc(nc);
    SUCEM_LAUNCH(k_pt_morton, grid_for(nc, TB), TB, 0, c->stream, nc, m->xyz.p, m->cells.p, lo[0],
                 lo[1], lo[2], sc[0], sc[1], sc[2], key.p, val.p);
    size_t bytes = 0;
    SUCEM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key.p, key_s.p, val.p,
                                        (uint32_t*)s->pt_cell.p, (size_t)nc, 0u, 30u, c->stream));
    DBuf<char> tmp(bytes);
    SUCEM_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, key.p, key_s.p, val.p,
                                        (uint32_t*)s->pt_cell.p, (size_t)nc, 0u, 30u, c->stream));
    SUCEM_LAUNCH(k_pt_rank, grid_for(nc, TB), TB, 0, c->stream, nc, s->pt_cell.p, rank.p);
    c->sync();
  }
  s->pt_face.alloc(std::max<int64_t>(nbf, 1));
  s->pt_fmeta.alloc(std::max<int64_t>(nbf, 1));
  if (nbf) {
    DBuf<uint32_t> key(nbf), key_s(nbf), val(nbf);
    SUCEM_LAUNCH(k_pt_face_keys, grid_for(nbf, TB), TB, 0, c->stream, nbf, s->bf_cell.p, rank.p,
                 key.p, val.p);
    const unsigned bits = (unsigned)key_bits(nc);
    size_t bytes = 0;
    SUCEM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key.p, key_s.p, val.p,
                                        (uint32_t*)s->pt_face.p, (size_t)nbf, 0u, bits, c->stream));
    DBuf<char> tmp(bytes);
    SUCEM_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, key.p, key_s.p, val.p,
                                        (uint32_t*)s->pt_face.p, (size_t)nbf, 0u, bits, c->stream));
    SUCEM_LAUNCH(k_pt_face_meta, grid_for(nbf, TB), TB, 0, c->stream, nbf, key_s.p, s->pt_face.p,
                 s->bf_local.p, s->pt_fmeta.p);
    c->sync();
  }
  SUCEM_HIP(hipDeviceGetAttribute(&s->pt_max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                  c->device));
  s->pt_state = 1;
}

// The maps of patches of cb cells (pm.cb = cb on return, or -cb when a patch would hold
// 65536 or more products).
void pt_build_maps(Context* c, Space* s, PatchMaps& pm, int cb) {
  const int64_t nc = s->mesh->nc, nbf = s->nbf;
  const int nloc = s->nloc, nfd = s->nfd;
  const int64_t np = (nc + cb - 1) / cb;
  const int64_t ne = nc * nloc + nbf * nfd;
  pm.fptr.alloc(np + 1);
  pm.fpos.alloc(std::max<int64_t>(nbf * nfd, 1));
  if (nbf) {
    // face ranks ascend in patch order (pt_fmeta >> 2)
    DBuf<uint32_t> fr(nbf);
    SUCEM_LAUNCH(k_pt_meta_rank, grid_for(nbf, TB), TB, 0, c->stream, nbf, s->pt_fmeta.p, fr.p);
    SUCEM_LAUNCH((k_pt_ptr<uint32_t>), grid_for(nbf, TB), TB, 0, c->stream, nbf, fr.p, 0,
                 (int64_t)cb, np, pm.fptr.p);
    c->sync();
  } else {
    SUCEM_HIP(hipMemsetAsync(pm.fptr.p, 0, (np + 1) * sizeof(int32_t), c->stream));
  }
  // product slots sorted by (patch, DOF, slot)
  DBuf<uint64_t> key(ne), key_s(ne);
  DBuf<uint32_t> val(ne), slot(ne);
  SUCEM_LAUNCH(k_pt_entries, grid_for(ne, TB), TB, 0, c->stream, nc, nloc, nbf, nfd, cb,
               s->pt_cell.p, s->cell_dofs.p, s->pt_fmeta.p, s->ref_face_dofs.p, key.p, val.p);
  {
    const unsigned bits = 32u + (unsigned)key_bits(np);
    size_t bytes = 0;
    SUCEM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key.p, key_s.p, val.p, slot.p, (size_t)ne,
                                        0u, bits, c->stream));
    DBuf<char> tmp(bytes);
    SUCEM_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, key.p, key_s.p, val.p, slot.p, (size_t)ne,
                                        0u, bits, c->stream));
    c->sync();
  }
  key.release();
  val.release();
  DBuf<int32_t> pstart(np + 1), flag(ne), uidx(ne);
  SUCEM_LAUNCH((k_pt_ptr<uint64_t>), grid_for(ne, TB), TB, 0, c->stream, ne, key_s.p, 32,
               (int64_t)1, np, pstart.p);
  SUCEM_LAUNCH(k_pt_heads, grid_for(ne, TB), TB, 0, c->stream, ne, key_s.p, flag.p);
  scan<int32_t>(c, flag.p, uidx.p, ne, true);
  const int64_t nu = read_i32(c, uidx.p + ne - 1);
  std::vector<int32_t> ps(np + 1);
  d2h(c, ps.data(), pstart.p, np + 1);
  c->sync();
  int64_t maxs = 0;
  for (int64_t p = 0; p < np; ++p) maxs = std::max<int64_t>(maxs, ps[p + 1] - ps[p]);
  if (maxs >= 65536) {   // the 16-bit places cannot address such a patch
    pm.cb = -cb;
    return;
  }
  pm.np = np;
  pm.nu = nu;
  pm.dof.alloc(nu);
  pm.seg.alloc(nu);
  pm.out.alloc(nu);
  pm.uptr.alloc(np + 1);
  pm.lp.alloc((size_t)nc * nloc);
  SUCEM_LAUNCH(k_pt_units, grid_for(ne, TB), TB, 0, c->stream, ne, key_s.p, uidx.p, pstart.p,
               pm.dof.p, pm.seg.p);
  SUCEM_LAUNCH(k_pt_uptr, grid_for(np + 1, TB), TB, 0, c->stream, np, ne, pstart.p, uidx.p,
               pm.uptr.p);
  SUCEM_LAUNCH(k_pt_places, grid_for(ne, TB), TB, 0, c->stream, ne, nc, nloc, key_s.p, slot.p,
               uidx.p, pstart.p, pm.uptr.p, pm.lp.p, pm.fpos.p);
  std::vector<int32_t> up(np + 1), fp(np + 1);
  d2h(c, up.data(), pm.uptr.p, np + 1);
  d2h(c, fp.data(), pm.fptr.p, np + 1);
  c->sync();
  int64_t lds = 0, maxfr = 0;
  for (int64_t p = 0; p < np; ++p) {
    lds = std::max<int64_t>(lds, std::max(ps[p + 1] - ps[p], up[p + 1] - up[p]));
    maxfr = std::max<int64_t>(maxfr, (int64_t)(fp[p + 1] - fp[p]) * nfd);
  }
  // shared DOFs: partial sums sorted by (DOF, patch)
  DBuf<int32_t> cnt(s->ndofs);
  SUCEM_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes(), c->stream));
  SUCEM_LAUNCH(k_pt_count, grid_for(nu, TB), TB, 0, c->stream, nu, pm.dof.p, cnt.p);
  DBuf<uint32_t> du_in(nu), dk(nu), du(nu);
  SUCEM_LAUNCH(k_pt_iota, grid_for(nu, TB), TB, 0, c->stream, nu, du_in.p);
  {
    const unsigned bits = (unsigned)key_bits(s->ndofs);
    size_t bytes = 0;
    SUCEM_HIP(rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t*)pm.dof.p, dk.p,
                                        du_in.p, du.p, (size_t)nu, 0u, bits, c->stream));
    DBuf<char> tmp(bytes);
    SUCEM_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, (const uint32_t*)pm.dof.p, dk.p,
                                        du_in.p, du.p, (size_t)nu, 0u, bits, c->stream));
    c->sync();
  }
  DBuf<int32_t> shared(nu), head(nu), ppos(nu), qi(nu);
  SUCEM_LAUNCH(k_pt_shared, grid_for(nu, TB), TB, 0, c->stream, nu, dk.p, cnt.p, shared.p,
               head.p);
  scan<int32_t>(c, shared.p, ppos.p, nu, false);
  scan<int32_t>(c, head.p, qi.p, nu, false);
  const int64_t npe = read_i32(c, ppos.p + nu - 1) + read_i32(c, shared.p + nu - 1);
  const int64_t npd = read_i32(c, qi.p + nu - 1) + read_i32(c, head.p + nu - 1);
  pm.qdof.alloc(std::max<int64_t>(npd, 1));
  pm.qptr.alloc(npd + 1);
  SUCEM_LAUNCH(k_pt_outs, grid_for(nu, TB), TB, 0, c->stream, nu, dk.p, du.p, shared.p, ppos.p,
               head.p, qi.p, pm.out.p, pm.qdof.p, pm.qptr.p);
  const int32_t npe32 = (int32_t)npe;
  SUCEM_HIP(hipMemcpyAsync(pm.qptr.p + npd, &npe32, sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream));
  c->sync();
  pm.npd = npd;
  pm.npe = npe;
  pm.lds = (int)lds;
  pm.maxfr = (int)maxfr;
  pm.cb = cb;
  pt_validate(c, s, pm, ps, up, fp);
  if (npe < (int64_t(1) << 27)) {
    pm.qcode.alloc(s->ndofs);
    DBuf<int32_t> too_many(1);
    SUCEM_LAUNCH(k_set_i32_pt, grid_for(s->ndofs, TB), TB, 0, c->stream, s->ndofs, pm.qcode.p, -1);
    SUCEM_LAUNCH(k_set_i32_pt, 1, 64, 0, c->stream, (int64_t)1, too_many.p, 0);
    if (npd)
      SUCEM_LAUNCH(k_pt_qcode, grid_for(npd, TB), TB, 0, c->stream, npd, pm.qptr.p, pm.qdof.p,
                   pm.qcode.p, too_many.p);
    if (read_i32(c, too_many.p)) pm.qcode.release();
  }
}

// Host check of every index the patch kernels follow (one pass over the maps when they
// are built): patch-DOF indices and product places inside the patch's LDS ranges,
// segments ascending and closing at the patch's product count, destinations inside the
// partial-sum buffer, faces inside their patch.  A map that fails is an internal error,
// reported before any kernel reads through it.
void pt_validate(Context* c, const Space* s, const PatchMaps& pm, const std::vector<int32_t>& ps,
                 const std::vector<int32_t>& up, const std::vector<int32_t>& fp) {
  const int64_t nc = s->mesh->nc, nbf = s->nbf, np = pm.np, nu = pm.nu;
  const int nloc = s->nloc, nfd = s->nfd, cb = pm.cb;
  std::vector<uint32_t> lp((size_t)nc * nloc);
  std::vector<int32_t> dof(nu), out(nu), meta(nbf), qptr(pm.npd + 1), qdof(pm.npd);
  std::vector<uint16_t> seg(nu), fpos((size_t)nbf * nfd);
  std::vector<int32_t> cell(nc);
  d2h(c, lp.data(), pm.lp.p, lp.size());
  d2h(c, dof.data(), pm.dof.p, nu);
  d2h(c, out.data(), pm.out.p, nu);
  d2h(c, seg.data(), pm.seg.p, nu);
  d2h(c, cell.data(), s->pt_cell.p, nc);
  d2h(c, meta.data(), s->pt_fmeta.p, nbf);
  d2h(c, fpos.data(), pm.fpos.p, fpos.size());
  d2h(c, qptr.data(), pm.qptr.p, qptr.size());
  d2h(c, qdof.data(), pm.qdof.p, qdof.size());
  c->sync();
  bool ok = ps[0] == 0 && up[0] == 0 && fp[0] == 0 && up[np] == nu && fp[np] == nbf &&
            ps[np] == nc * nloc + nbf * nfd;
  std::vector<char> seen(nc, 0);
  for (int64_t r = 0; ok && r < nc; ++r) {
    ok = cell[r] >= 0 && cell[r] < nc && !seen[cell[r]];
    if (ok) seen[cell[r]] = 1;
  }
  for (int64_t p = 0; ok && p < np; ++p) {
    const int64_t S = ps[p + 1] - ps[p], U = up[p + 1] - up[p];
    ok = S > 0 && U > 0 && S <= pm.lds && U <= pm.lds && fp[p + 1] >= fp[p];
    for (int64_t r = p * cb; ok && r < std::min<int64_t>(nc, (p + 1) * cb); ++r)
      for (int l = 0; ok && l < nloc; ++l) {
        const uint32_t v = lp[(size_t)l * nc + r];
        ok = (int64_t)(v & 0xffffu) < U && (int64_t)(v >> 16) < S;
      }
    int64_t prev = 0;
    for (int64_t u = up[p]; ok && u < up[p + 1]; ++u) {
      ok = seg[u] > prev && dof[u] >= 0 && dof[u] < s->ndofs && out[u] >= -1 && out[u] < pm.npe;
      prev = seg[u];
    }
    ok = ok && prev == S;
    for (int64_t f = fp[p]; ok && f < fp[p + 1]; ++f) {
      ok = (meta[f] >> 2) / cb == p;
      for (int a = 0; ok && a < nfd; ++a) ok = fpos[(size_t)f * nfd + a] < S;
    }
  }
  for (int64_t i = 0; ok && i < pm.npd; ++i)
    ok = qptr[i] < qptr[i + 1] && qdof[i] >= 0 && qdof[i] < s->ndofs;
  ok = ok && (pm.npd == 0 || (qptr[0] == 0 && qptr[pm.npd] == pm.npe));
  SUCEM_CHECK(ok, SUCEM_ERR_INTERNAL, "patch operator maps failed their consistency check");
}

template <typename R, int ORDER, int NFD, int CB>
void launch_patch(Context* c, const System* sys, const PatchMaps& pm, Coef k,
                  const uint8_t* mask, const void* x, void* y, const void* w, int epi,
                  const RedSite& rs, int Gp, int Gt) {
  using C = typename C2<R>::T;
  const Space* s = sys->space;
  const R* coef;
  const R* hf;
  if constexpr (sizeof(R) == 4) {
    coef = sys->pt_coef32.p; hf = sys->pt_face32.p;
  } else {
    coef = sys->pt_coef64.p; hf = sys->pt_face64.p;
  }
  const R c0 = (sys->which & SUCEM_MAT_S0) ? (R)k.s0 : R(0);
  const size_t lds = (size_t)(pm.lds + pm.maxfr) * sizeof(C);
#define PT_ARGS pm.np, Gp, Gt, s->mesh->nc, s->nbf, pm.lds, coef, pm.lp.p, pm.uptr.p, pm.dof.p, \
                pm.seg.p, pm.out.p, pm.fptr.p, s->pt_fmeta.p, pm.fpos.p, hf, s->ref_face_dofs.p, \
                s->ref_face.p, (R)k.s, (R)k.t, c0, mask, (const C*)x, (C*)y, (C*)sys->pt_part.p, \
                (const C*)w, rs
  if (epi == 0) SUCEM_LAUNCH((k_mf_patch<R, ORDER, NFD, 0, CB>), Gp, CB, lds, c->stream, PT_ARGS);
  else if (epi == 1) SUCEM_LAUNCH((k_mf_patch<R, ORDER, NFD, 1, CB>), Gp, CB, lds, c->stream, PT_ARGS);
  else if (epi == 2) SUCEM_LAUNCH((k_mf_patch<R, ORDER, NFD, 2, CB>), Gp, CB, lds, c->stream, PT_ARGS);
  else if (epi == 3) SUCEM_LAUNCH((k_mf_patch<R, ORDER, NFD, 3, CB>), Gp, CB, lds, c->stream, PT_ARGS);
  else SUCEM_LAUNCH((k_mf_patch<R, ORDER, NFD, 4, CB>), Gp, CB, lds, c->stream, PT_ARGS);
#undef PT_ARGS
}

}  // namespace

bool pt_prepare(Context* c, System* sys, int cb) {
  Space* s = sys->space;
  const int slot = cb_slot(cb);
  SUCEM_CHECK(slot >= 0, SUCEM_ERR_INVALID, "cells per patch must be 128, 192 or 256");
  {
    std::lock_guard<std::mutex> hold(s->lazy);
    if (s->pt_state == 0) pt_build_order(c, s);
    if (s->pt_state > 0 && s->pt[slot].cb == 0)
      pt_build_maps(c, s, s->pt[slot], cb);
  }
  if (s->pt_state < 0 || s->pt[slot].cb != cb) return false;
  if (!sys->pt_ready) {
    mf_prepare(c, sys);   // the per-cell coefficients, permuted below
    const int64_t nc = s->mesh->nc, nbf = s->nbf;
    sys->pt_coef64.alloc((size_t)12 * nc);
    sys->pt_coef32.alloc((size_t)12 * nc);
    SUCEM_LAUNCH((k_pt_permute<double>), grid_for(12 * nc, TB), TB, 0, c->stream, nc, 12,
                 s->pt_cell.p, sys->mf_coef64.p, sys->pt_coef64.p);
    SUCEM_LAUNCH((k_pt_permute<float>), grid_for(12 * nc, TB), TB, 0, c->stream, nc, 12,
                 s->pt_cell.p, sys->mf_coef32.p, sys->pt_coef32.p);
    sys->pt_face64.alloc(std::max<int64_t>(6 * nbf, 1));
    sys->pt_face32.alloc(std::max<int64_t>(6 * nbf, 1));
    if (nbf) {
      SUCEM_LAUNCH((k_pt_permute<double>), grid_for(6 * nbf, TB), TB, 0, c->stream, nbf, 6,
                   s->pt_face.p, sys->mf_face64.p, sys->pt_face64.p);
      SUCEM_LAUNCH((k_pt_permute<float>), grid_for(6 * nbf, TB), TB, 0, c->stream, nbf, 6,
                   s->pt_face.p, sys->mf_face32.p, sys->pt_face32.p);
    }
    c->sync();
    sys->pt_ready = true;
  }
  if ((int64_t)sys->pt_part.n < std::max<int64_t>(s->pt[slot].npe, 1)) {
    c->sync();
    sys->pt_part.alloc(std::max<int64_t>(s->pt[slot].npe, 1));
  }
  return true;
}

template <typename R>
bool pt_fits(const Space* s, int cb) {
  const int slot = cb_slot(cb);
  if (slot < 0 || s->pt_state <= 0 || s->pt[slot].cb != cb) return false;
  const PatchMaps& pm = s->pt[slot];
  return (int64_t)(pm.lds + pm.maxfr) * (int64_t)sizeof(typename C2<R>::T) <=
         (int64_t)s->pt_max_lds;
}
template bool pt_fits<float>(const Space*, int);
template bool pt_fits<double>(const Space*, int);

SpmvCfg pt_cfg(const Space* s, int cb, int G) {
  SpmvCfg f;
  f.mf = 2;
  f.U = cb;
  f.G = G;
  const int slot = cb_slot(cb);
  f.P = (int)((s->pt[slot < 0 ? 0 : slot].np + 7) / 8 * 8) + G;
  return f;
}

template <typename R>
void pt_apply(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x, void* y,
              const void* w, int epi, const RedSite& rs, const SpmvCfg& f) {
  using C = typename C2<R>::T;
  const Space* s = sys->space;
  const int slot = cb_slot(f.U);
  SUCEM_CHECK(sys->pt_ready && slot >= 0 && s->pt[slot].cb == f.U, SUCEM_ERR_INVALID,
              "patch operator not prepared");
  const PatchMaps& pm = s->pt[slot];
  const int Gp = (int)((pm.np + 7) / 8 * 8);
  const int Gg = f.G;
  const int Gt = Gp + Gg;
  SUCEM_CHECK(Gg > 0 && Gg % 8 == 0 && f.parts() == Gt, SUCEM_ERR_INVALID,
              "bad patch operator launch configuration");
  SUCEM_CHECK(pt_fits<R>(s, f.U), SUCEM_ERR_UNSUPPORTED,
              "the patch operator needs more LDS per workgroup than the device has");
  SUCEM_CHECK((int64_t)sys->pt_part.n >= pm.npe, SUCEM_ERR_INVALID, "patch operator not prepared");
  if (s->order == 2) {
    if (f.U == 128) launch_patch<R, 2, 8, 128>(c, sys, pm, k, mask, x, y, w, epi, rs, Gp, Gt);
    else if (f.U == 192) launch_patch<R, 2, 8, 192>(c, sys, pm, k, mask, x, y, w, epi, rs, Gp, Gt);
    else launch_patch<R, 2, 8, 256>(c, sys, pm, k, mask, x, y, w, epi, rs, Gp, Gt);
  } else {
    SUCEM_CHECK(f.U == 256, SUCEM_ERR_UNSUPPORTED, "order-1 patches hold 256 cells");
    launch_patch<R, 1, 3, 256>(c, sys, pm, k, mask, x, y, w, epi, rs, Gp, Gt);
  }
  const int npd = (int)pm.npd;
#define PS_ARGS npd, Gg, Gt, Gp, pm.qptr.p, pm.qdof.p, (const C*)sys->pt_part.p, mask, (C*)y, \
                (const C*)w, rs
  if (epi == 0) SUCEM_LAUNCH((k_mf_psum<R, 0>), Gg, TB, 0, c->stream, PS_ARGS);
  else if (epi == 1) SUCEM_LAUNCH((k_mf_psum<R, 1>), Gg, TB, 0, c->stream, PS_ARGS);
  else SUCEM_LAUNCH((k_mf_psum<R, 2>), Gg, TB, 0, c->stream, PS_ARGS);
#undef PS_ARGS
}
template void pt_apply<float>(Context*, const System*, Coef, const uint8_t*, const void*, void*,
                              const void*, int, const RedSite&, const SpmvCfg&);

int pt_deferred_parts(const System* sys, const SpmvCfg& f) {
  if (f.mf != 2 || !sys->pt_ready) return 0;
  const int slot = cb_slot(f.U);
  if (slot < 0) return 0;
  const PatchMaps& pm = sys->space->pt[slot];
  if (pm.cb != f.U || !pm.qcode.p) return 0;
  return (int)((pm.np + 7) / 8 * 8);
}

template <typename R>
void pt_apply_deferred(Context* c, const System* sys, Coef k, const uint8_t* mask, const void* x,
                       void* y, const void* w, const RedSite& rs, const SpmvCfg& f) {
  const int Gp = pt_deferred_parts(sys, f);
  SUCEM_CHECK(Gp > 0, SUCEM_ERR_INVALID, "deferred patch operator not available");
  SUCEM_CHECK(pt_fits<R>(sys->space, f.U), SUCEM_ERR_UNSUPPORTED,
              "the patch operator needs more LDS per workgroup than the device has");
  const PatchMaps& pm = sys->space->pt[cb_slot(f.U)];
  const int epi = rs.paired ? 4 : 3;
  if (sys->space->order == 2) {
    if (f.U == 128) launch_patch<R, 2, 8, 128>(c, sys, pm, k, mask, x, y, w, epi, rs, Gp, Gp);
    else if (f.U == 192) launch_patch<R, 2, 8, 192>(c, sys, pm, k, mask, x, y, w, epi, rs, Gp, Gp);
    else launch_patch<R, 2, 8, 256>(c, sys, pm, k, mask, x, y, w, epi, rs, Gp, Gp);
  } else {
    launch_patch<R, 1, 3, 256>(c, sys, pm, k, mask, x, y, w, epi, rs, Gp, Gp);
  }
}
template void pt_apply_deferred<float>(Context*, const System*, Coef, const uint8_t*,
                                       const void*, void*, const void*, const RedSite&,
                                       const SpmvCfg&);
template void pt_apply_deferred<double>(Context*, const System*, Coef, const uint8_t*,
                                        const void*, void*, const void*, const RedSite&,
                                        const SpmvCfg&);

const int32_t* pt_qcode(const System* sys, const SpmvCfg& f) {
  return sys->space->pt[cb_slot(f.U)].qcode.p;
}
const void* pt_partials(const System* sys) { return sys->pt_part.p; }
template void pt_apply<double>(Context*, const System*, Coef, const uint8_t*, const void*, void*,
                               const void*, int, const RedSite&, const SpmvCfg&);

}  // namespace sucem
