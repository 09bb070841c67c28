// Smoothing-prior operators per geometry (init.f90:295-541), on the device.
//
// The node Laplacian `luni` has at most 9 entries per row, at column offsets
// (dk, dl) in {-1,0,1}^2 inside the node grid of the plane.  k_luni builds each row
// in a 9-slot register accumulator using the exact sequence of additions of
// init.f90:331-394, so the values are bit-identical to the dense reference matrix.
// From the slots:
//   k_lell   ELL(9) of L (rows 2*nnode) and of L^T (rows 2*ndof), gen_sparse_lmat
//            semantics (drop |v| <= 1e-8, pad with index 1 / value 0)
//   k_llcols sparse columns of L^T L (<= 25 entries, 5x5 node window), summed in
//            ascending node order like the reference's dense product
#include "smcf_device.h"
#include "smcf_internal.h"

namespace smcf {

SMCF_DEV int slot_of(int dk, int dl) { return (dk + 1) + 3 * (dl + 1); }

__global__ void k_luni(MeshDev m, const double* __restrict__ geo, double* __restrict__ luni, int B,
                       const int* __restrict__ sel) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  int b = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  if (n >= m.nnode) return;
  int ip = m.node_plane[n];
  const double* g = geo + ((size_t)b * m.nplane + ip) * 16;
  double dxi = g[9], deta = g[10];
  int nxi = m.plane_nxi[ip], neta = m.plane_neta[ip];
  int k = m.node_k[n], l = m.node_l[n];
  double v[9];
#pragma unroll
  for (int s = 0; s < 9; ++s) v[s] = 0.0;
  double cx = 1.0 / (dxi * dxi), ce = 1.0 / (deta * deta), cc = 1.0 / (4.0 * dxi * deta);
  v[slot_of(0, 0)] = v[slot_of(0, 0)] - 2.0 / (dxi * dxi);
  if (k == nxi + 1) v[slot_of(-1, 0)] += cx; else v[slot_of(1, 0)] += cx;
  if (k == 1) v[slot_of(1, 0)] += cx; else v[slot_of(-1, 0)] += cx;
  if (l == neta + 1) v[slot_of(0, -1)] += ce; else v[slot_of(0, 1)] += ce;
  if (l == 1) v[slot_of(0, 1)] += ce; else v[slot_of(0, -1)] += ce;
  int W = nxi + 1, L = neta + 1;
  if (k == W && l == L) v[slot_of(-1, -1)] += cc;
  else if (k == W) v[slot_of(-1, 1)] += cc;
  else if (l == L) v[slot_of(1, -1)] += cc;
  else v[slot_of(1, 1)] += cc;
  if (k == 1 && l == L) v[slot_of(1, -1)] -= cc;
  else if (k == 1) v[slot_of(1, 1)] -= cc;
  else if (l == L) v[slot_of(-1, -1)] -= cc;
  else v[slot_of(-1, 1)] -= cc;
  if (k == 1 && l == 1) v[slot_of(1, 1)] += cc;
  else if (k == 1) v[slot_of(1, -1)] += cc;
  else if (l == 1) v[slot_of(-1, 1)] += cc;
  else v[slot_of(-1, -1)] += cc;
  if (k == W && l == 1) v[slot_of(-1, 1)] -= cc;
  else if (k == W) v[slot_of(-1, -1)] -= cc;
  else if (l == 1) v[slot_of(1, 1)] -= cc;
  else v[slot_of(1, -1)] -= cc;
  double* o = luni + ((size_t)b * m.nnode + n) * 9;
#pragma unroll
  for (int s = 0; s < 9; ++s) o[s] = v[s];
}

// node index of the slot neighbour, or -1 when outside the plane's grid
SMCF_DEV int slot_node(const MeshDev& m, int n, int dk, int dl) {
  int ip = m.node_plane[n];
  int k = m.node_k[n] + dk, l = m.node_l[n] + dl;
  int nxi = m.plane_nxi[ip], neta = m.plane_neta[ip];
  if (k < 1 || k > nxi + 1 || l < 1 || l > neta + 1) return -1;
  return m.plane_node0[ip] + (k - 1) + (l - 1) * (nxi + 1);
}

__global__ void k_lell(MeshDev m, const double* __restrict__ luni, int* __restrict__ li, double* __restrict__ lv,
                       int* __restrict__ lti, double* __restrict__ ltv, int B, const int* __restrict__ sel) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int b = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  const double* lu = luni + (size_t)b * m.nnode * 9;
  if (r < 2 * m.nnode) {
    // row r = 2*n + c of L: columns 2*dof(j) + c for DOF nodes j in ascending order
    int n = r >> 1, c = r & 1, cnt = 0;
    int* oi = li + ((size_t)b * 2 * m.nnode + r) * 9;
    double* ov = lv + ((size_t)b * 2 * m.nnode + r) * 9;
    for (int dl = -1; dl <= 1; ++dl)
      for (int dk = -1; dk <= 1; ++dk) {
        int j = slot_node(m, n, dk, dl);
        if (j < 0) continue;
        int d = m.dof_of_node[j];
        double v = lu[(size_t)n * 9 + slot_of(dk, dl)];
        if (d >= 0 && fabs(v) > 1e-8 && cnt < 9) {
          oi[cnt] = 2 * d + c + 1;
          ov[cnt] = v;
          cnt++;
        }
      }
    for (; cnt < 9; ++cnt) { oi[cnt] = 1; ov[cnt] = 0.0; }
  } else if (r < 2 * m.nnode + 2 * m.ndof) {
    // row of L^T for column 2*d + c: rows 2*n + c with luni(n, j) != 0, ascending n
    int q = r - 2 * m.nnode;
    int d = q >> 1, c = q & 1, cnt = 0;
    int j = m.id_dof[d] - 1;
    int* oi = lti + ((size_t)b * 2 * m.ndof + q) * 9;
    double* ov = ltv + ((size_t)b * 2 * m.ndof + q) * 9;
    for (int dl = 1; dl >= -1; --dl)
      for (int dk = 1; dk >= -1; --dk) {
        int n = slot_node(m, j, -dk, -dl);  // n + (dk, dl) == j
        if (n < 0) continue;
        double v = lu[(size_t)n * 9 + slot_of(dk, dl)];
        if (fabs(v) > 1e-8 && cnt < 9) {
          oi[cnt] = 2 * n + c + 1;
          ov[cnt] = v;
          cnt++;
        }
      }
    for (; cnt < 9; ++cnt) { oi[cnt] = 1; ov[cnt] = 0.0; }
  }
}

// Column i = 2*d + c of L^T L: LL(r, i) = sum_n luni(n, j_r) luni(n, j_i) over nodes n in
// ascending order (all nodes, same component), kept in a 5x5 window around node j_i.
__global__ void k_llcols(MeshDev m, const double* __restrict__ luni, int* __restrict__ idx,
                         double* __restrict__ val, int B, const int* __restrict__ sel) {
  // separate multiply and add, as in the reference's dense product
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  if (i >= m.ndim) return;
  const double* lu = luni + (size_t)b * m.nnode * 9;
  int d = i >> 1, c = i & 1;
  int ji = m.id_dof[d] - 1;
  double acc[25];
#pragma unroll
  for (int w = 0; w < 25; ++w) acc[w] = 0.0;
  // nodes n with luni(n, ji) possibly nonzero: n = ji - (dk, dl); ascending n
  for (int dl = 1; dl >= -1; --dl)
    for (int dk = 1; dk >= -1; --dk) {
      int n = slot_node(m, ji, -dk, -dl);
      if (n < 0) continue;
      double a = lu[(size_t)n * 9 + slot_of(dk, dl)];
      // partner columns j_r = n + (ek, el); offset from ji = (ek - dk, el - dl)
      for (int el = -1; el <= 1; ++el)
        for (int ek = -1; ek <= 1; ++ek) {
          int jr = slot_node(m, n, ek, el);
          if (jr < 0) continue;
          double bv = lu[(size_t)n * 9 + slot_of(ek, el)];
          int w = (ek - dk + 2) + 5 * (el - dl + 2);
          acc[w] = acc[w] + bv * a;
        }
    }
  int* oi = idx + ((size_t)b * m.ndim + i) * 25;
  double* ov = val + ((size_t)b * m.ndim + i) * 25;
  int cnt = 0;
  int ip = m.node_plane[ji];
  int W = m.plane_nxi[ip] + 1;
  // ascending row order: window (wl, wk) lexicographic gives ascending node ids
  for (int wl = -2; wl <= 2; ++wl)
    for (int wk = -2; wk <= 2; ++wk) {
      int kk = m.node_k[ji] + wk, ll = m.node_l[ji] + wl;
      if (kk < 1 || kk > W || ll < 1 || ll > m.plane_neta[ip] + 1) continue;
      int jr = m.plane_node0[ip] + (kk - 1) + (ll - 1) * W;
      int dr = m.dof_of_node[jr];
      if (dr < 0) continue;
      oi[cnt] = 2 * dr + c;
      ov[cnt] = acc[(wk + 2) + 5 * (wl + 2)];
      cnt++;
    }
  for (; cnt < 25; ++cnt) { oi[cnt] = -1; ov[cnt] = 0.0; }
}

hipError_t launch_laplacian(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel) {
  hipLaunchKernelGGL(k_luni, dim3((m.nnode + 127) / 128, B), dim3(128), 0, s, m, bt.geo, bt.luni, B, sel);
  int rows = 2 * m.nnode + 2 * m.ndof;
  hipLaunchKernelGGL(k_lell, dim3((rows + 127) / 128, B), dim3(128), 0, s, m, bt.luni, bt.li, bt.lv, bt.lti,
                     bt.ltv, B, sel);
  hipLaunchKernelGGL(k_llcols, dim3((m.ndim + 127) / 128, B), dim3(128), 0, s, m, bt.luni, bt.llc_idx,
                     bt.llc_val, B, sel);
  return hipGetLastError();
}

}  // namespace smcf
