// k_geom.h -- grid metrics on the device: ghost nodes (src/grid.f90:137-236),
// face area vectors, areas and unit normals (src/geometry.f90:50-314), cell
// volumes as the larger of two five-tetrahedron splits (src/geometry.f90:
// 342-496), centroids (src/geometry.f90:500-545) and the brute-force wall
// distance (src/wall/wall_dist.f90:84-131).
#pragma once
// ghost cells and geometry keep the reference's unfused operation order
#pragma clang fp contract(off)
#include "blk.h"

namespace fest {

// linear extrapolation of three ghost node layers along direction d, over
// the whole padded node plane across it; passes run i, then j, then k
__global__ void k_ghost_nodes(Blk b, int d) {
  int a = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 2;
  int bb = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 2;
  int c = blockIdx.z;  // component
  int na = d == 0 ? b.jmx : b.imx;
  int nb = d == 2 ? b.jmx : b.kmx;
  if (a > na + 3 || bb > nb + 3) return;
  double *x = b.nodes + (long)c * (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  int N = d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx;
  auto at = [&](int n) -> double & {
    if (d == 0) return x[ndi(b, n, a, bb)];
    if (d == 1) return x[ndi(b, a, n, bb)];
    return x[ndi(b, a, bb, n)];
  };
  at(0) = 2 * at(1) - at(2);
  at(-1) = 2 * at(0) - at(1);
  at(-2) = 2 * at(-1) - at(0);
  at(N + 1) = 2 * at(N) - at(N - 1);
  at(N + 2) = 2 * at(N + 1) - at(N);
  at(N + 3) = 2 * at(N + 2) - at(N + 1);
}

// area vector, area and unit normal of every face of direction d
__global__ void k_face_metrics(Blk b, int d) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 2;
  int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 2;
  int k = (int)blockIdx.z - 2;
  int ie = b.imx + 2 + (d == 0), je = b.jmx + 2 + (d == 1), ke = b.kmx + 2 + (d == 2);
  if (i > ie || j > je || k > ke) return;
  long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  const double *X = b.nodes, *Y = X + nn, *Z = Y + nn;
  long p0 = ndi(b, i, j, k), p1, p2, p3;
  // d1 = P(p1) - P(p0), d2 = P(p2) - P(p3)
  if (d == 0) { p1 = ndi(b, i, j + 1, k + 1); p2 = ndi(b, i, j, k + 1); p3 = ndi(b, i, j + 1, k); }
  else if (d == 1) { p1 = ndi(b, i + 1, j, k + 1); p2 = ndi(b, i + 1, j, k); p3 = ndi(b, i, j, k + 1); }
  else { p1 = ndi(b, i + 1, j + 1, k); p2 = ndi(b, i, j + 1, k); p3 = ndi(b, i + 1, j, k); }
  double d1x = X[p1] - X[p0], d1y = Y[p1] - Y[p0], d1z = Z[p1] - Z[p0];
  double d2x = X[p2] - X[p3], d2y = Y[p2] - Y[p3], d2z = Z[p2] - Z[p3];
  double nx = 0.5 * (d1y * d2z - d1z * d2y);
  double ny = 0.5 * (d1z * d2x - d1x * d2z);
  double nz = 0.5 * (d1x * d2y - d1y * d2x);
  double A = sqrt(sqd(nx) + sqd(ny) + sqd(nz));
  if (A != 0.) { nx = nx / A; ny = ny / A; nz = nz / A; }
  long f = fi(b, d, i, j, k);
  b.fA[d][f] = A; b.fx[d][f] = nx; b.fy[d][f] = ny; b.fz[d][f] = nz;
}

// pole face (id -7), src/geometry.f90:101-186 and 227-232: the pole face and
// its three ghost faces (layers 1..-2, or N..N+3 at a max face) get zero area
// and the unit normal of the first interior face (2, or N-1), over the whole
// padded plane.  blockIdx.z is the layer; the source layer is never written.
__global__ void k_pole_faces(Blk b, int face) {
  const int d = face >> 1, hi = face & 1;
  int a = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 2;
  int c = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 2;
  int ae = (d == 0 ? b.jmx : b.imx) + 2, ce = (d == 2 ? b.jmx : b.kmx) + 2;
  if (a > ae || c > ce) return;
  int N = d == 0 ? b.imx : d == 1 ? b.jmx : b.kmx;
  int l = (int)blockIdx.z;
  int n = hi ? N + l : 1 - l, src = hi ? N - 1 : 2;
  long t, o;
  if (d == 0) { t = fi(b, 0, n, a, c); o = fi(b, 0, src, a, c); }
  else if (d == 1) { t = fi(b, 1, a, n, c); o = fi(b, 1, a, src, c); }
  else { t = fi(b, 2, a, c, n); o = fi(b, 2, a, c, src); }
  b.fA[d][t] = 0.;
  b.fx[d][t] = b.fx[d][o]; b.fy[d][t] = b.fy[d][o]; b.fz[d][t] = b.fz[d][o];
}

__device__ double vol_tet(const double *p1, const double *p2, const double *p3, const double *p4) {
  double v = (p4[0] - p1[0]) * ((p2[1] - p1[1]) * (p3[2] - p1[2]) - (p2[2] - p1[2]) * (p3[1] - p1[1]))
           + (p4[1] - p1[1]) * ((p2[2] - p1[2]) * (p3[0] - p1[0]) - (p2[0] - p1[0]) * (p3[2] - p1[2]))
           + (p4[2] - p1[2]) * ((p2[0] - p1[0]) * (p3[1] - p1[1]) - (p2[1] - p1[1]) * (p3[0] - p1[0]));
  return -v / 6.;
}

// volumes (cells 0..imx; 1.0 elsewhere) and centroids (all padded cells)
__global__ void k_cells(Blk b) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 2;
  int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 2;
  int k = (int)blockIdx.z - 2;
  if (i > b.imx + 2 || j > b.jmx + 2 || k > b.kmx + 2) return;
  long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  const double *X = b.nodes, *Y = X + nn, *Z = Y + nn;
  const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  double p[8][3];
  for (int q = 0; q < 8; ++q) {
    long n = ndi(b, i + o[q][0], j + o[q][1], k + o[q][2]);
    p[q][0] = X[n]; p[q][1] = Y[n]; p[q][2] = Z[n];
  }
  long c = ci(b, i, j, k);
  double vol = 1.;
  if (i >= 0 && i <= b.imx && j >= 0 && j <= b.jmx && k >= 0 && k <= b.kmx) {
    double v1 = 0.;
    v1 = v1 + vol_tet(p[0], p[4], p[7], p[5]);
    v1 = v1 + vol_tet(p[6], p[7], p[5], p[2]);
    v1 = v1 + vol_tet(p[7], p[3], p[0], p[2]);
    v1 = v1 + vol_tet(p[5], p[0], p[2], p[7]);
    v1 = v1 + vol_tet(p[0], p[1], p[5], p[2]);
    double v = 0.;
    v = v + vol_tet(p[1], p[5], p[4], p[6]);
    v = v + vol_tet(p[7], p[4], p[6], p[3]);
    v = v + vol_tet(p[4], p[0], p[1], p[3]);
    v = v + vol_tet(p[6], p[1], p[3], p[4]);
    v = v + vol_tet(p[1], p[2], p[6], p[3]);
    vol = fmx(v, v1);
    if (vol <= 0.0) atomicOr(b.err, ERR_NONPHYS);
  }
  b.vol[c] = vol;
  // centroid: the eight nodes in src/geometry.f90:511-519 order
  const int co[8] = {0, 1, 2, 6, 5, 3, 7, 4};
  for (int a = 0; a < 3; ++a) {
    double s = p[co[0]][a];
    for (int q = 1; q < 8; ++q) s = s + p[co[q]][a];
    (a == 0 ? b.cx : a == 1 ? b.cy : b.cz)[c] = 0.125 * s;
  }
}

// nearest wall-node distance of every node: min over the surface list of the
// squared distance, then one square root (monotone, so equal to the minimum
// of the square roots the reference takes)
__global__ void __launch_bounds__(256) k_node_dist(Blk b, const double *wall, int nw, double *nd) {
  __shared__ double sw[3][256];
  long nn = (long)(b.imx + 6) * (b.jmx + 6) * (b.kmx + 6);
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const double *X = b.nodes, *Y = X + nn, *Z = Y + nn;
  double x = 0, y = 0, z = 0;
  if (t < nn) { x = X[t]; y = Y[t]; z = Z[t]; }
  double best = 1.e+40;
  for (int base = 0; base < nw; base += 256) {
    int m = base + threadIdx.x;
    __syncthreads();
    if (m < nw) { sw[0][threadIdx.x] = wall[3 * m]; sw[1][threadIdx.x] = wall[3 * m + 1]; sw[2][threadIdx.x] = wall[3 * m + 2]; }
    __syncthreads();
    int cnt = min(256, nw - base);
    for (int q = 0; q < cnt; ++q) {
      double dx = sw[0][q] - x, dy = sw[1][q] - y, dz = sw[2][q] - z;
      double d2 = dx * dx + dy * dy + dz * dz;
      best = fmn(best, d2);
    }
  }
  if (t < nn) nd[t] = fmn(1.e+20, sqrt(best));
}

// cell distance: average of the eight node distances in src/wall/
// wall_dist.f90:119-126 order
__global__ void k_cell_dist(Blk b, const double *nd) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 2;
  int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) - 2;
  int k = (int)blockIdx.z - 2;
  if (i > b.imx + 2 || j > b.jmx + 2 || k > b.kmx + 2) return;
  b.dist[ci(b, i, j, k)] = 0.125 * (nd[ndi(b, i, j, k)] + nd[ndi(b, i, j + 1, k)] + nd[ndi(b, i, j + 1, k + 1)] +
                                    nd[ndi(b, i, j, k + 1)] + nd[ndi(b, i + 1, j, k + 1)] + nd[ndi(b, i + 1, j, k)] +
                                    nd[ndi(b, i + 1, j + 1, k)] + nd[ndi(b, i + 1, j + 1, k + 1)]);
}

// lctm2015's wall-normal sensor of each interior cell (src/CC.f90 and
// src/source.f90:404-407): D = the Green-Gauss gradient of the wall
// distance (CC.f90 compute_gradient; find_DCCVn passes the distance, not
// the velocity component, CC.f90:117-120), n = D / (|D| + 1e-12) as the
// set-up normalises it (CC.f90:89-97), dvdy = D . n.  The distance is fixed
// once the grid is, so this runs at set-up, not per residual.
__global__ void k_lctm_sensor(Blk b) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  int k = (int)blockIdx.z + 1;
  if (i > b.imx - 1 || j > b.jmx - 1 || k > b.kmx - 1) return;
  const double *var = b.dist;
  const long c = ci(b, i, j, k);
  const double vc = var[c];
  double D[3];
  for (int d = 0; d < 3; ++d) {
    const double *in = b.fx[0], *jn = b.fx[1], *kn = b.fx[2];
    if (d == 1) { in = b.fy[0]; jn = b.fy[1]; kn = b.fy[2]; }
    if (d == 2) { in = b.fz[0]; jn = b.fz[1]; kn = b.fz[2]; }
    const long f0 = fi(b, 0, i, j, k), f1 = fi(b, 0, i + 1, j, k);
    const long g0 = fi(b, 1, i, j, k), g1 = fi(b, 1, i, j + 1, k);
    const long h0 = fi(b, 2, i, j, k), h1 = fi(b, 2, i, j, k + 1);
    D[d] = (-(var[ci(b, i - 1, j, k)] + vc) * in[f0] * b.fA[0][f0]
            - (var[ci(b, i, j - 1, k)] + vc) * jn[g0] * b.fA[1][g0]
            - (var[ci(b, i, j, k - 1)] + vc) * kn[h0] * b.fA[2][h0]
            + (var[ci(b, i + 1, j, k)] + vc) * in[f1] * b.fA[0][f1]
            + (var[ci(b, i, j + 1, k)] + vc) * jn[g1] * b.fA[1][g1]
            + (var[ci(b, i, j, k + 1)] + vc) * kn[h1] * b.fA[2][h1]) /
           (2 * b.vol[c]);
  }
  const double mag = sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
  const double nx = D[0] / (mag + 1e-12), ny = D[1] / (mag + 1e-12), nz = D[2] / (mag + 1e-12);
  b.dvdy[c] = D[0] * nx + D[1] * ny + D[2] * nz;
}

__global__ void k_fill(double *p, long n, double v) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

}  // namespace fest
