/*
 * Synthetic cylinder meshes for the benchmark workloads (host code, built into
 * libcrimson_mesh.so; no GPU).  An nx x ny x nz hex grid over [-R,R]^2 x
 * [0,L]; hexes are kept when their centroid radius is <= R (or all of them
 * when clip == 0); each kept hex is split into the 6 Kuhn tets along its
 * (0,0,0)-(1,1,1) diagonal, which is conforming across hexes.
 *
 * Numbering: used nodes in lexicographic (i, j, k) order with i fastest;
 * hexes k-slowest, then kept columns with i fastest; the 6 tets of a hex in
 * the order of the permutations (x,y,z),(x,z,y),(y,x,z),(y,z,x),(z,x,y),
 * (z,y,x) of the unit steps walked from the hex's low corner.  Tets are
 * oriented so that det[x1-x4, x2-x4, x3-x4] > 0 (first two nodes swapped
 * otherwise), the orientation with positive WdetJ in the reference's
 * e3metric (e3metric.f90:42-67).
 */
#include <stdlib.h>
#include <string.h>

static const int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

static int kept(int nx, int ny, double R, int clip, int i, int j) {
  if (i < 0 || j < 0 || i >= nx || j >= ny) return 0;
  if (!clip) return 1;
  const double hx = 2.0 * R / nx, hy = 2.0 * R / ny;
  const double ci = (i + 0.5) * hx - R, cj = (j + 0.5) * hy - R;
  return ci * ci + cj * cj <= R * R;
}

/* counts of the mesh: elements, nodes, outlet triangles */
int cyl_mesh_counts(int nx, int ny, int nz, double R, int clip, long *nel, long *numnp, long *ntri) {
  if (nx < 1 || ny < 1 || nz < 1 || !(R > 0)) return 1;
  long ncol = 0, nused = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) ncol += kept(nx, ny, R, clip, i, j);
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i)
      nused += kept(nx, ny, R, clip, i - 1, j - 1) | kept(nx, ny, R, clip, i, j - 1) |
               kept(nx, ny, R, clip, i - 1, j) | kept(nx, ny, R, clip, i, j);
  *nel = 6 * ncol * nz;
  *numnp = nused * (nz + 1);
  *ntri = 2 * ncol;
  return 0;
}

/* x(numnp,3) column-major; ien(nel,4) column-major 1-based; tag(numnp):
 * bit 0 inlet (z = 0), bit 1 outlet (z = L), bit 2 wall (touches a removed
 * or outside column); tri(ntri,3) row-major 0-based outlet triangles with
 * normal +z. */
int cyl_mesh_fill(int nx, int ny, int nz, double R, double L, int clip, double *x, int *ien,
                  unsigned char *tag, int *tri) {
  long nel, numnp, ntri;
  if (cyl_mesh_counts(nx, ny, nz, R, clip, &nel, &numnp, &ntri)) return 1;
  const int NX = nx + 1, NY = ny + 1;
  const double hx = 2.0 * R / nx, hy = 2.0 * R / ny, hz = L / nz;
  int *rank = malloc(sizeof(int) * NX * NY);
  unsigned char *wall2 = malloc((size_t)NX * NY);
  int nused = 0;
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const int a = kept(nx, ny, R, clip, i - 1, j - 1), b = kept(nx, ny, R, clip, i, j - 1),
                c = kept(nx, ny, R, clip, i - 1, j), d = kept(nx, ny, R, clip, i, j);
      rank[i + NX * j] = (a | b | c | d) ? nused++ : -1;
      wall2[i + NX * j] = !(a & b & c & d);
    }
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const int r = rank[i + NX * j];
        if (r < 0) continue;
        const long n = (long)k * nused + r;
        x[n] = i * hx - R;
        x[n + numnp] = j * hy - R;
        x[n + 2 * numnp] = k * hz;
        tag[n] = (unsigned char)((k == 0 ? 1 : 0) | (k == nz ? 2 : 0) | (wall2[i + NX * j] ? 4 : 0));
      }
  long e = 0, t = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        if (!kept(nx, ny, R, clip, i, j)) continue;
        for (int p = 0; p < 6; ++p) {
          int o[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}};
          o[1][PERM[p][0]] = 1;
          o[2][PERM[p][0]] = 1;
          o[2][PERM[p][1]] = 1;
          long v[4];
          for (int a = 0; a < 4; ++a)
            v[a] = (long)(k + o[a][2]) * nused + rank[(i + o[a][0]) + NX * (j + o[a][1])];
          double d[3][3];
          for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) d[a][c] = x[v[a] + c * numnp] - x[v[3] + c * numnp];
          const double det = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
                             d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
                             d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
          if (det < 0) {
            const long s = v[0];
            v[0] = v[1];
            v[1] = s;
          }
          for (int a = 0; a < 4; ++a) ien[e + a * nel] = (int)(v[a] + 1);
          if (k == nz - 1 && PERM[p][0] == 2) {
            /* the three vertices other than the low corner lie on z = L */
            long f[3], m = 0;
            for (int a = 0; a < 4; ++a)
              if (x[v[a] + 2 * numnp] > L - 0.5 * hz) f[m++] = v[a];
            const double ax = x[f[1]] - x[f[0]], ay = x[f[1] + numnp] - x[f[0] + numnp];
            const double bx = x[f[2]] - x[f[0]], by = x[f[2] + numnp] - x[f[0] + numnp];
            if (ax * by - ay * bx < 0) {
              const long s = f[1];
              f[1] = f[2];
              f[2] = s;
            }
            tri[3 * t + 0] = (int)f[0];
            tri[3 * t + 1] = (int)f[1];
            tri[3 * t + 2] = (int)f[2];
            ++t;
          }
          ++e;
        }
      }
  free(rank);
  free(wall2);
  return (e == nel && t == ntri) ? 0 : 2;
}
