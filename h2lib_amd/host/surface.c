/* Surfaces: structured refinement of macro triangulations and the
 * per-triangle geometry (normals, Gram determinants, mesh widths).
 *
 * The refinement numbers vertices and triangles exactly as the reference's
 * build_from_macrosurface3d_surface3d (Library/macrosurface3d.c:707-850):
 * macro vertices and the interior points of macro edges in order of first
 * use, then the interior points of each macro triangle, and the triangles
 * row by row.  The macro data of the sphere and the cube (vertex
 * coordinates, triangles, edge orientations) are those of
 * Library/macrosurface3d.c:81-200 and 258-388; the edge opposite each
 * triangle vertex is found from the oriented edge list.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "h2bem.h"

typedef struct {
  uint32_t nv, ne, nt;
  const double (*x)[3];
  const uint32_t (*e)[2];
  const uint32_t (*t)[3];
  int project; /* 1: project onto the unit sphere, 2: onto the cylinder */
} macro;

static const double sphere_x[6][3] = {
    {0.0, 0.0, 1.0}, {1.0, 0.0, 0.0},  {0.0, 0.0, -1.0},
    {-1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, -1.0, 0.0}};
static const uint32_t sphere_e[12][2] = {
    {0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 4},
    {2, 4}, {3, 4}, {0, 5}, {1, 5}, {2, 5}, {3, 5}};
static const uint32_t sphere_t[8][3] = {{0, 1, 4}, {1, 2, 4}, {2, 3, 4},
                                        {3, 0, 4}, {0, 5, 1}, {1, 5, 2},
                                        {2, 5, 3}, {3, 5, 0}};

static const double cube_x[8][3] = {
    {-1.0, -1.0, -1.0}, {-1.0, 1.0, -1.0}, {-1.0, -1.0, 1.0},
    {-1.0, 1.0, 1.0},   {1.0, -1.0, -1.0}, {1.0, 1.0, -1.0},
    {1.0, -1.0, 1.0},   {1.0, 1.0, 1.0}};
static const uint32_t cube_e[18][2] = {
    {1, 0}, {0, 3}, {3, 1}, {2, 0}, {3, 2}, {5, 7}, {7, 4}, {4, 5}, {4, 6},
    {6, 7}, {2, 6}, {6, 3}, {3, 7}, {5, 1}, {1, 4}, {4, 0}, {1, 7}, {0, 6}};
static const uint32_t cube_t[12][3] = {
    {1, 0, 3}, {3, 0, 2}, {4, 5, 7}, {4, 7, 6}, {3, 2, 6}, {3, 6, 7},
    {4, 0, 1}, {4, 1, 5}, {0, 6, 2}, {4, 6, 0}, {1, 3, 7}, {1, 7, 5}};

/* The cylinder of radius 1 along the x axis from -4 to 4, closed by two
 * discs (Library/macrosurface3d.c:431-700): 22 vertices on five circles,
 * 60 edges, 40 triangles. */
#define HALF_SQRT M_SQRT1_2 /* sqrt(0.5), correctly rounded */
static const double cyl_x[22][3] = {
    {-4.0, 0.0, 1.0}, {-4.0, 1.0, 0.0}, {-4.0, 0.0, -1.0}, {-4.0, -1.0, 0.0},
    {-4.0, 0.0, 0.0},
    {-2.0, HALF_SQRT, HALF_SQRT}, {-2.0, HALF_SQRT, -HALF_SQRT},
    {-2.0, -HALF_SQRT, -HALF_SQRT}, {-2.0, -HALF_SQRT, HALF_SQRT},
    {0.0, 0.0, 1.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, -1.0}, {0.0, -1.0, 0.0},
    {2.0, HALF_SQRT, HALF_SQRT}, {2.0, HALF_SQRT, -HALF_SQRT},
    {2.0, -HALF_SQRT, -HALF_SQRT}, {2.0, -HALF_SQRT, HALF_SQRT},
    {4.0, 0.0, 1.0}, {4.0, 1.0, 0.0}, {4.0, 0.0, -1.0}, {4.0, -1.0, 0.0},
    {4.0, 0.0, 0.0}};
static const uint32_t cyl_e[60][2] = {
    {0, 1},   {1, 4},   {4, 0},   {4, 2},   {2, 1},   {2, 3},   {3, 4},
    {0, 3},   {0, 5},   {0, 8},   {1, 5},   {1, 6},   {2, 6},   {2, 7},
    {3, 7},   {3, 8},   {5, 9},   {5, 10},  {6, 10},  {6, 11},  {7, 11},
    {7, 12},  {8, 9},   {8, 12},  {0, 9},   {1, 10},  {2, 11},  {3, 12},
    {17, 9},  {18, 10}, {19, 11}, {20, 12}, {13, 9},  {13, 10}, {14, 10},
    {14, 11}, {15, 11}, {15, 12}, {16, 9},  {16, 12}, {17, 13}, {17, 16},
    {18, 13}, {18, 14}, {19, 14}, {19, 15}, {20, 15}, {20, 16}, {17, 18},
    {18, 21}, {21, 17}, {21, 19}, {19, 18}, {19, 20}, {20, 21}, {17, 20},
    {5, 13},  {6, 14},  {7, 15},  {8, 16}};
static const uint32_t cyl_t[40][3] = {
    {0, 1, 4},    {1, 2, 4},    {2, 3, 4},    {0, 4, 3},    {0, 5, 1},
    {1, 6, 2},    {2, 7, 3},    {3, 8, 0},    {0, 9, 5},    {0, 8, 9},
    {1, 5, 10},   {1, 10, 6},   {2, 6, 11},   {2, 11, 7},   {3, 7, 12},
    {3, 12, 8},   {5, 9, 13},   {5, 13, 10},  {6, 10, 14},  {6, 14, 11},
    {7, 11, 15},  {7, 15, 12},  {8, 12, 16},  {8, 16, 9},   {9, 17, 13},
    {9, 16, 17},  {10, 13, 18}, {10, 18, 14}, {11, 14, 19}, {11, 19, 15},
    {12, 15, 20}, {12, 20, 16}, {13, 17, 18}, {14, 18, 19}, {15, 19, 20},
    {16, 20, 17}, {17, 21, 18}, {18, 21, 19}, {19, 21, 20}, {17, 20, 21}};

/* Point of macro triangle i at local coordinates (a, b). */
static void
macro_point(const macro *m, uint32_t i, double a, double b, double *out)
{
  const double *p0 = m->x[m->t[i][0]];
  const double *p1 = m->x[m->t[i][1]];
  const double *p2 = m->x[m->t[i][2]];
  for (int d = 0; d < 3; d++)
    out[d] = p0[d] * (1.0 - a - b) + p1[d] * a + p2[d] * b;
  if (m->project == 1) {
    double r = sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
    out[0] /= r;
    out[1] /= r;
    out[2] /= r;
  } else if (m->project == 2) {
    /* cylinder_parametrization, Library/macrosurface3d.c:394-428: the
     * mantle is projected onto the unit circle in (y, z); on the end discs
     * the L1 ball is mapped onto the L2 ball */
    const double n2 = sqrt(out[1] * out[1] + out[2] * out[2]);
    if (fabs(out[0]) != 4.0) {
      out[1] /= n2;
      out[2] /= n2;
    } else {
      const double n1 = fabs(out[1]) + fabs(out[2]);
      if (n2 > 0.0) {
        out[1] = out[1] * n1 / n2;
        out[2] = out[2] * n1 / n2;
      }
    }
  }
}

/* Edge of macro triangle i opposite its vertex j. */
static uint32_t
macro_edge(const macro *m, uint32_t i, int j)
{
  uint32_t a = m->t[i][(j + 1) % 3], b = m->t[i][(j + 2) % 3];
  for (uint32_t k = 0; k < m->ne; k++)
    if ((m->e[k][0] == a && m->e[k][1] == b) ||
        (m->e[k][0] == b && m->e[k][1] == a))
      return k;
  abort();
}

static void
prepare(h2b_surface *s)
{
  double hmin = 1e30, hmax = 0.0;
  for (uint32_t i = 0; i < s->triangles; i++) {
    const double *p0 = s->x + 3 * s->t[3 * i];
    const double *p1 = s->x + 3 * s->t[3 * i + 1];
    const double *p2 = s->x + 3 * s->t[3 * i + 2];
    double u[3], v[3], w[3], c[3];
    for (int d = 0; d < 3; d++) {
      u[d] = p1[d] - p0[d];
      v[d] = p2[d] - p0[d];
      w[d] = p2[d] - p1[d];
    }
    c[0] = u[1] * v[2] - u[2] * v[1];
    c[1] = u[2] * v[0] - u[0] * v[2];
    c[2] = u[0] * v[1] - u[1] * v[0];
    double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    s->g[i] = len;
    for (int d = 0; d < 3; d++)
      s->n[3 * i + d] = c[d] / len;
    /* height of the triangle relative to its shortest edge, as
     * Library/surface3d.c:88-101 computes it */
    double a = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    double b = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double cc = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double sh = a < b ? a : b;
    sh = sh < cc ? sh : cc;
    double height = 2.0 * sh;
    a *= a;
    b *= b;
    cc *= cc;
    height = sqrt(2 * (a * b + b * cc + a * cc) - (a * a + b * b + cc * cc))
      / height;
    hmin = height < hmin ? height : hmin;
    hmax = height > hmax ? height : hmax;
  }
  s->hmin = hmin;
  s->hmax = hmax;
}

static h2b_surface *
alloc_surface(uint32_t nv, uint32_t nt)
{
  h2b_surface *s = (h2b_surface *) calloc(1, sizeof(h2b_surface));
  s->vertices = nv;
  s->triangles = nt;
  s->x = (double *) malloc(sizeof(double) * 3 * nv);
  s->t = (uint32_t *) malloc(sizeof(uint32_t) * 3 * nt);
  s->n = (double *) malloc(sizeof(double) * 3 * nt);
  s->g = (double *) malloc(sizeof(double) * nt);
  return s;
}

static h2b_surface *
refine(const macro *m, uint32_t split)
{
  if (split < 1)
    return NULL;
  const uint32_t S = split, W = split + 1;
  uint32_t nv = m->nv + m->ne * (S - 1) + m->nt * (S - 1) * (S - 2) / 2;
  if (S == 1)
    nv = m->nv;
  uint32_t nt = m->nt * S * S;
  h2b_surface *s = alloc_surface(nv, nt);
  uint32_t *vv = (uint32_t *) malloc(sizeof(uint32_t) * m->nv);
  uint32_t *ve = (uint32_t *) malloc(sizeof(uint32_t) * m->ne);
  char *vdone = (char *) calloc(m->nv, 1), *edone = (char *) calloc(m->ne, 1);
  uint32_t *vloc = (uint32_t *) malloc(sizeof(uint32_t) * W * W);
  uint32_t (*ms)[3] = malloc(sizeof(uint32_t[3]) * m->nt);
  uint32_t vc = 0, tc = 0;

  for (uint32_t i = 0; i < m->nt; i++)
    for (int j = 0; j < 3; j++)
      ms[i][j] = macro_edge(m, i, j);

  /* vertices of macro vertices and macro edges, in order of first use */
  for (uint32_t i = 0; i < m->nt; i++) {
    for (int j = 0; j < 3; j++) {
      uint32_t v = m->t[i][j];
      if (vdone[v])
        continue;
      vv[v] = vc;
      macro_point(m, i, j == 1 ? 1.0 : 0.0, j == 2 ? 1.0 : 0.0,
                  s->x + 3 * vc);
      vc++;
      vdone[v] = 1;
    }
    for (int j = 0; j < 3; j++) {
      uint32_t e = ms[i][j];
      if (edone[e])
        continue;
      ve[e] = vc;
      for (uint32_t k = 1; k < S; k++) {
        double al = (m->e[e][0] == m->t[i][(j + 1) % 3])
          ? (double) k / S : (double) (S - k) / S;
        double a = j == 0 ? 1.0 - al : (j == 1 ? 0.0 : al);
        double b = j == 0 ? al : (j == 1 ? 1.0 - al : 0.0);
        macro_point(m, i, a, b, s->x + 3 * vc);
        vc++;
      }
      edone[e] = 1;
    }
  }

  /* interior vertices and triangles of each macro triangle; vloc[a + b W]
   * numbers the lattice point with barycentric steps a (towards t1) and
   * b (towards t2) */
  for (uint32_t i = 0; i < m->nt; i++) {
    const uint32_t t0 = m->t[i][0], t1 = m->t[i][1], t2 = m->t[i][2];
    vloc[0] = vv[t0];
    vloc[S] = vv[t1];
    vloc[S * W] = vv[t2];
    uint32_t e0 = ms[i][0], e1 = ms[i][1], e2 = ms[i][2];
    for (uint32_t j = 1; j < S; j++) {
      /* edge t1 -> t2 */
      if (m->e[e0][0] == t1)
        vloc[(S - j) + j * W] = ve[e0] + j - 1;
      else
        vloc[j + (S - j) * W] = ve[e0] + j - 1;
      /* edge t2 -> t0 */
      if (m->e[e1][0] == t2)
        vloc[(S - j) * W] = ve[e1] + j - 1;
      else
        vloc[j * W] = ve[e1] + j - 1;
      /* edge t0 -> t1 */
      if (m->e[e2][0] == t0)
        vloc[j] = ve[e2] + j - 1;
      else
        vloc[S - j] = ve[e2] + j - 1;
    }
    for (uint32_t j = 1; j < S; j++)
      for (uint32_t k = 1; k + j < S; k++) {
        vloc[k + j * W] = vc;
        macro_point(m, i, (double) k / S, (double) j / S, s->x + 3 * vc);
        vc++;
      }
    for (uint32_t j = 0; j < S; j++) {
      uint32_t k;
      for (k = 0; k + j + 1 < S; k++) {
        uint32_t *tt = s->t + 3 * tc;
        tt[0] = vloc[k + j * W];
        tt[1] = vloc[(k + 1) + j * W];
        tt[2] = vloc[k + (j + 1) * W];
        tt += 3;
        tt[0] = vloc[(k + 1) + (j + 1) * W];
        tt[1] = vloc[k + (j + 1) * W];
        tt[2] = vloc[(k + 1) + j * W];
        tc += 2;
      }
      uint32_t *tt = s->t + 3 * tc;
      tt[0] = vloc[k + j * W];
      tt[1] = vloc[(k + 1) + j * W];
      tt[2] = vloc[k + (j + 1) * W];
      tc++;
    }
  }
  free(vv);
  free(ve);
  free(vdone);
  free(edone);
  free(vloc);
  free(ms);
  if (vc != nv || tc != nt) {
    h2b_surface_free(s);
    return NULL;
  }
  prepare(s);
  return s;
}

h2b_surface *
h2b_sphere_surface(uint32_t split)
{
  macro m = {6, 12, 8, sphere_x, sphere_e, sphere_t, 1};
  return refine(&m, split);
}

h2b_surface *
h2b_cube_surface(uint32_t split)
{
  macro m = {8, 18, 12, cube_x, cube_e, cube_t, 0};
  return refine(&m, split);
}

h2b_surface *
h2b_cylinder_surface(uint32_t split)
{
  macro m = {22, 60, 40, cyl_x, cyl_e, cyl_t, 2};
  return refine(&m, split);
}

h2b_surface *
h2b_surface_from_arrays(uint32_t vertices, const double *x,
                        uint32_t triangles, const uint32_t *t)
{
  for (size_t i = 0; i < 3 * (size_t) triangles; i++)
    if (t[i] >= vertices)
      return NULL;
  h2b_surface *s = alloc_surface(vertices, triangles);
  memcpy(s->x, x, sizeof(double) * 3 * vertices);
  memcpy(s->t, t, sizeof(uint32_t) * 3 * triangles);
  prepare(s);
  return s;
}

void
h2b_surface_transform(h2b_surface *s, const double *R, double scale,
                      const double *shift)
{
  for (uint32_t i = 0; i < s->vertices; i++) {
    double *p = s->x + 3 * i, q[3];
    for (int r = 0; r < 3; r++)
      q[r] = scale * (R[3 * r] * p[0] + R[3 * r + 1] * p[1]
                      + R[3 * r + 2] * p[2]) + (shift ? shift[r] : 0.0);
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
  }
  prepare(s);
}

void
h2b_surface_free(h2b_surface *s)
{
  if (!s)
    return;
  free(s->x);
  free(s->t);
  free(s->n);
  free(s->g);
  free(s);
}
