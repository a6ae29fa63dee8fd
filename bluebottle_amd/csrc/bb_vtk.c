/*
 * bb_vtk.c -- ParaView output in the reference's XML VTK files
 * (src/vtk.c:26-322, 935-3589):
 *   out.pvd             collection; every output appends three DataSets
 *                       (grid, particles, quadrature nodes) at ttime
 *   out_N.pvtr          the parallel header of the rectilinear grid
 *   out_N_0.vtr         cell data p (%e), phase (%d) and the face
 *                       velocities averaged to cell centres (%e), ASCII
 *   out_N.vtp           one point per particle and per periodic image of a
 *                       particle that straddles a face, with its index,
 *                       radius, velocities, accelerations, body axes and
 *                       forces (total F adds the buoyant weight)
 *   out_nodes_N.vtp     the 26 Lebedev nodes of every point of out_N.vtp
 *                       at r = rs, with their interference codes
 *   out_turb.pvd, out_turb_N.pvtr, out_turb_N_0.vtr
 *                       the precursor domain's output (src/vtk.c:324-611):
 *                       one DataSet per output, cell data p and vel only
 *   out_ghost.pvd, out_ghost_N.pvtr, out_ghost_N_0.vtr
 *                       the debugging output with the ghost cells
 *                       (src/vtk.c:612-933), plus out_N.vtp
 * The images follow the reference's order: for each particle, the single
 * face images (x low, x high, y low, y high, z low, z high), the edge
 * images of the xy, zy and xz pairs, the corner images, then the particle
 * itself; a particle within r of a low face is drawn shifted by +L, of a
 * high face by -L, whether or not the axis is periodic.  Here the images
 * come from one table instead of one block per image, and the image count
 * is the number actually drawn.
 *
 * Departure: the reference takes each node sphere's radius from parts[i]
 * with i the image index (out of range once images exist); here it is
 * the imaged particle's own rs, the same whenever all particles share rs.
 * The node wrap in z compares against the x end, as there.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/bluebottle.h"
#include "bb_host.h"

extern int rec_paraview_stepnum_out, rec_prec_stepnum_out;

#define VTK_PI 3.14159265358979323846

/* one image: the side of each axis it is taken across (-1 low face, +1
 * high face, 0 none), in the reference's order */
static const signed char IMAGES[26][3] = {
  {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
  /* x with y */
  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}, {1, 1, 0},
  /* z with y */
  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1},
  /* x with z */
  {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {1, 0, 1},
  /* corners */
  {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1},
  {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1},
};

/* does particle P lie within its radius of the face on `side` of axis a */
static int near_face(const part_struct *P, int a, int side)
{
  const real c[3] = {P->x, P->y, P->z};
  const real lo[3] = {Dom.xs, Dom.ys, Dom.zs};
  const real hi[3] = {Dom.xe, Dom.ye, Dom.ze};
  if(side < 0) return c[a] < lo[a] + P->r;
  return c[a] > hi[a] - P->r;
}

/* the particles as drawn: copies with shifted centres and, in ind[], the
 * index of the particle each came from.  Returns the count (or -1). */
static int virtual_parts(part_struct **out, int **ind_out)
{
  const real L[3] = {Dom.xl, Dom.yl, Dom.zl};
  int cap = nparts * 27, n = 0;
  part_struct *pv = (part_struct *) malloc((size_t) (cap ? cap : 1)
    * sizeof(part_struct));
  int *ind = (int *) malloc((size_t) (cap ? cap : 1) * sizeof(int));
  if(pv == NULL || ind == NULL) { free(pv); free(ind); return -1; }
  for(int i = 0; i < nparts; i++) {
    const part_struct *P = &parts[i];
    for(int m = 0; m < 26; m++) {
      int take = 1;
      real shift[3] = {0., 0., 0.};
      for(int a = 0; a < 3 && take; a++) {
        const int s = IMAGES[m][a];
        if(s == 0) continue;
        take = near_face(P, a, s);
        shift[a] = s < 0 ? L[a] : -L[a];
      }
      if(!take) continue;
      pv[n] = *P;
      pv[n].x = P->x + shift[0];
      pv[n].y = P->y + shift[1];
      pv[n].z = P->z + shift[2];
      ind[n++] = i;
    }
    pv[n] = *P;
    ind[n++] = i;
  }
  *out = pv;
  *ind_out = ind;
  return n;
}

static FILE *open_out(const char *dir, const char *fmt, int step)
{
  char name[64], path[2 * FILE_NAME_SIZE + 64];
  snprintf(name, sizeof(name), fmt, step);
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "w");
  if(f == NULL) fprintf(stderr, "Could not open file %s\n", path);
  return f;
}

/* ---------------------------------------------------------------------
 * Collections: out.pvd (main domain), out_turb.pvd (precursor domain)
 * ------------------------------------------------------------------- */
static int pvd_init(const char *dir, const char *name)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "w");
  if(f == NULL) {
    fprintf(stderr, "Could not open file %s\n", path);
    return -1;
  }
  fprintf(f, "<VTKFile type=\"Collection\">\n<Collection>\n");
  fprintf(f, "</Collection>\n</VTKFile>");
  fclose(f);
  return 0;
}

int bb_vtk_init(const char *dir) { return pvd_init(dir, "out.pvd"); }
int bb_vtk_turb_init(const char *dir) { return pvd_init(dir, "out_turb.pvd"); }

/* append one DataSet per entry of kind[] (printf patterns of the step) */
static int pvd_append(const char *dir, const char *name, int step,
  const char *const *kind, int nkind)
{
  char path[2 * FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "r+");
  if(f == NULL) {
    if(pvd_init(dir, name)) return -1;
    f = fopen(path, "r+");
    if(f == NULL) return -1;
  }
  /* overwrite the 24-byte footer "</Collection>\n</VTKFile>" */
  fseek(f, -24, SEEK_END);
  for(int part = 0; part < nkind; part++) {
    char file[64];
    snprintf(file, sizeof(file), kind[part], step);
    fprintf(f, "<DataSet timestep=\"%e\" part=\"%d\" file=\"%s\"/>\n", ttime,
      part, file);
  }
  fprintf(f, "</Collection>\n</VTKFile>");
  fclose(f);
  return 0;
}

/* ---------------------------------------------------------------------
 * Grid: PREFIX_N.pvtr + PREFIX_N_0.vtr (one subdomain).  The main domain
 * (src/vtk.c:26-322) writes p, phase and vel and rounds the piece extent
 * of the header; the precursor domain (src/vtk.c:324-611) writes p and vel
 * and truncates it.  The piece file truncates in both.
 * ------------------------------------------------------------------- */
static int grid_files(const char *dir, const char *prefix, int step,
  int with_phase, int round_header)
{
  const grid_info *gc = &Dom.Gcc;
  const dom_struct *sd = &dom[0];
  char fmt[64];
  snprintf(fmt, sizeof(fmt), "%s_%%d.pvtr", prefix);
  FILE *f = open_out(dir, fmt, step);
  if(f == NULL) return -1;
  const real ext[6] = {(sd->xe - sd->xs) / Dom.dx, (sd->xs - Dom.xs) / Dom.dx,
    (sd->ye - sd->ys) / Dom.dy, (sd->ys - Dom.ys) / Dom.dy,
    (sd->ze - sd->zs) / Dom.dz, (sd->zs - Dom.zs) / Dom.dz};
  int e[6];
  for(int q = 0; q < 6; q++)
    e[q] = round_header ? (int) round(ext[q]) : (int) ext[q];
  const char *scalars = with_phase ? "p phase" : "p";
  fprintf(f, "<VTKFile type=\"PRectilinearGrid\">\n");
  fprintf(f, "<PRectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\" "
    "GhostLevel=\"0\">\n", Dom.xn, Dom.yn, Dom.zn);
  fprintf(f, "<PCellData Scalars=\"%s\" Vectors=\"vel\">\n", scalars);
  fprintf(f, "<PDataArray type=\"Float32\" Name=\"p\"/>\n");
  if(with_phase) fprintf(f, "<PDataArray type=\"Int32\" Name=\"phase\"/>\n");
  fprintf(f, "<PDataArray type=\"Float32\" Name=\"vel\" "
    "NumberOfComponents=\"3\"/>\n");
  fprintf(f, "</PCellData>\n<PCoordinates>\n");
  for(int a = 0; a < 3; a++)
    fprintf(f, "<PDataArray type=\"Float32\" Name=\"%c\"/>\n", "xyz"[a]);
  fprintf(f, "</PCoordinates>\n");
  fprintf(f, "<Piece Extent=\"%d %d %d %d %d %d\" Source=\"%s_%d_0.vtr\"/>\n",
    e[1], e[1] + e[0], e[3], e[3] + e[2], e[5], e[5] + e[4], prefix, step);
  fprintf(f, "</PRectilinearGrid>\n</VTKFile>\n");
  fclose(f);

  const int xl = (int) ext[0], xs0 = (int) ext[1];
  const int yl = (int) ext[2], ys0 = (int) ext[3];
  const int zl = (int) ext[4], zs0 = (int) ext[5];
  snprintf(fmt, sizeof(fmt), "%s_%%d_0.vtr", prefix);
  f = open_out(dir, fmt, step);
  if(f == NULL) return -1;
  fprintf(f, "<VTKFile type=\"RectilinearGrid\">\n");
  fprintf(f, "<RectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\" "
    "GhostLevel=\"0\">\n", Dom.xn, Dom.yn, Dom.zn);
  fprintf(f, "<Piece Extent=\"%d %d %d %d %d %d\">\n", xs0, xs0 + xl, ys0,
    ys0 + yl, zs0, zs0 + zl);
  fprintf(f, "<CellData Scalars=\"%s\" Vectors=\"vel\">\n", scalars);
  const int k0 = zs0 + gc->ks, k1 = zs0 + zl + gc->ks;
  const int j0 = ys0 + gc->js, j1 = ys0 + yl + gc->js;
  const int i0 = xs0 + gc->is, i1 = xs0 + xl + gc->is;
  fprintf(f, "<DataArray type=\"Float32\" Name=\"p\">\n");
  for(int k = k0; k < k1; k++)
    for(int j = j0; j < j1; j++)
      for(int i = i0; i < i1; i++)
        fprintf(f, "%e ", p[i + j * gc->s1b + k * gc->s2b]);
  fprintf(f, "\n</DataArray>\n");
  if(with_phase) {
    fprintf(f, "<DataArray type=\"Int32\" Name=\"phase\">\n");
    for(int k = k0; k < k1; k++)
      for(int j = j0; j < j1; j++)
        for(int i = i0; i < i1; i++)
          fprintf(f, "%d ", phase[i + j * gc->s1b + k * gc->s2b]);
    fprintf(f, "\n</DataArray>\n");
  }
  /* face velocities averaged to the cell centre, computed as written */
  fprintf(f, "<DataArray type=\"Float32\" Name=\"vel\" NumberOfComponents=\"3\""
    " format=\"ascii\">\n");
  const grid_info *gx = &Dom.Gfx, *gy = &Dom.Gfy, *gz = &Dom.Gfz;
  for(int k = k0; k < k1; k++)
    for(int j = j0; j < j1; j++)
      for(int i = i0; i < i1; i++) {
        const int cx = i + j * gx->s1b + k * gx->s2b;
        const int cy = i + j * gy->s1b + k * gy->s2b;
        const int cz = i + j * gz->s1b + k * gz->s2b;
        fprintf(f, "%e %e %e ", 0.5 * (u[cx] + u[cx + 1]),
          0.5 * (v[cy] + v[cy + gy->s1b]), 0.5 * (w[cz] + w[cz + gz->s2b]));
      }
  fprintf(f, "\n</DataArray>\n</CellData>\n<Coordinates>\n");
  const real s0[3] = {sd->xs, sd->ys, sd->zs};
  const real h[3] = {Dom.dx, Dom.dy, Dom.dz};
  const int n[3] = {xl, yl, zl};
  for(int a = 0; a < 3; a++) {
    fprintf(f, "<DataArray type=\"Float32\" Name=\"%c\">\n", "xyz"[a]);
    for(int q = 0; q <= n[a]; q++) fprintf(f, "%lf ", s0[a] + h[a] * q);
    fprintf(f, "\n</DataArray>\n");
  }
  fprintf(f, "</Coordinates>\n</Piece>\n</RectilinearGrid>\n</VTKFile>\n");
  fclose(f);
  return 0;
}

int bb_vtk_grid(const char *dir, int step)
{
  return grid_files(dir, "out", step, 1, 1);
}

int bb_vtk_turb_grid(const char *dir, int step)
{
  return grid_files(dir, "out_turb", step, 0, 0);
}

/* ---------------------------------------------------------------------
 * The debugging grid output with the ghost cells (src/vtk.c:612-933):
 * out_ghost_N.pvtr + out_ghost_N_0.vtr over all (xn+2)(yn+2)(zn+2) cells,
 * with p, phase, phase_shell, and the face values of u, u* and the face
 * flags averaged to the cell centres; the coordinates start one cell
 * below the domain.  The piece extent is the truncated cell count plus
 * the two ghost layers.
 * ------------------------------------------------------------------- */
/* cells [i0, i1) x [j0, j1) x [k0, k1) of the ghosted grid */
typedef struct { int i0, i1, j0, j1, k0, k1; } cell_range;

static void centre3(FILE *f, const real *x, const real *y, const real *z,
  cell_range r)
{
  const grid_info *gx = &Dom.Gfx, *gy = &Dom.Gfy, *gz = &Dom.Gfz;
  for(int k = r.k0; k < r.k1; k++)
    for(int j = r.j0; j < r.j1; j++)
      for(int i = r.i0; i < r.i1; i++) {
        const int cx = i + j * gx->s1b + k * gx->s2b;
        const int cy = i + j * gy->s1b + k * gy->s2b;
        const int cz = i + j * gz->s1b + k * gz->s2b;
        fprintf(f, "%e %e %e ", 0.5 * (x[cx] + x[cx + 1]),
          0.5 * (y[cy] + y[cy + gy->s1b]), 0.5 * (z[cz] + z[cz + gz->s2b]));
      }
}

static void centre3i(FILE *f, const int *x, const int *y, const int *z,
  cell_range r)
{
  const grid_info *gx = &Dom.Gfx, *gy = &Dom.Gfy, *gz = &Dom.Gfz;
  for(int k = r.k0; k < r.k1; k++)
    for(int j = r.j0; j < r.j1; j++)
      for(int i = r.i0; i < r.i1; i++) {
        const int cx = i + j * gx->s1b + k * gx->s2b;
        const int cy = i + j * gy->s1b + k * gy->s2b;
        const int cz = i + j * gz->s1b + k * gz->s2b;
        fprintf(f, "%lf %lf %lf ", 0.5 * (x[cx] + x[cx + 1]),
          0.5 * (y[cy] + y[cy + gy->s1b]), 0.5 * (z[cz] + z[cz + gz->s2b]));
      }
}

int bb_vtk_ghost_grid(const char *dir, int step)
{
  const grid_info *gc = &Dom.Gcc;
  const dom_struct *sd = &dom[0];
  const int xl = (int) ((sd->xe - sd->xs) / Dom.dx + 2 * DOM_BUF);
  const int xs0 = (int) ((sd->xs - Dom.xs) / Dom.dx);
  const int yl = (int) ((sd->ye - sd->ys) / Dom.dy + 2 * DOM_BUF);
  const int ys0 = (int) ((sd->ys - Dom.ys) / Dom.dy);
  const int zl = (int) ((sd->ze - sd->zs) / Dom.dz + 2 * DOM_BUF);
  const int zs0 = (int) ((sd->zs - Dom.zs) / Dom.dz);
  const int wx = Dom.xn + 2 * DOM_BUF, wy = Dom.yn + 2 * DOM_BUF;
  const int wz = Dom.zn + 2 * DOM_BUF;
  const char *vectors[3] = {"vel", "vel_star", "flag"};
  FILE *f = open_out(dir, "out_ghost_%d.pvtr", step);
  if(f == NULL) return -1;
  fprintf(f, "<VTKFile type=\"PRectilinearGrid\">\n");
  fprintf(f, "<PRectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\" "
    "GhostLevel=\"0\">\n", wx, wy, wz);
  fprintf(f, "<PCellData Scalars=\"p phase phase_shell\" "
    "Vectors=\"vel vel_star flag\">\n");
  fprintf(f, "<PDataArray type=\"Float32\" Name=\"p\"/>\n");
  fprintf(f, "<PDataArray type=\"Int32\" Name=\"phase\"/>\n");
  fprintf(f, "<PDataArray type=\"Int32\" Name=\"phase_shell\"/>\n");
  for(int q = 0; q < 3; q++)
    fprintf(f, "<PDataArray type=\"Float32\" Name=\"%s\" "
      "NumberOfComponents=\"3\"/>\n", vectors[q]);
  fprintf(f, "</PCellData>\n<PCoordinates>\n");
  for(int a = 0; a < 3; a++)
    fprintf(f, "<PDataArray type=\"Float32\" Name=\"%c\"/>\n", "xyz"[a]);
  fprintf(f, "</PCoordinates>\n");
  fprintf(f, "<Piece Extent=\"%d %d %d %d %d %d\" "
    "Source=\"out_ghost_%d_0.vtr\"/>\n", xs0, xs0 + xl, ys0, ys0 + yl, zs0,
    zs0 + zl, step);
  fprintf(f, "</PRectilinearGrid>\n</VTKFile>\n");
  fclose(f);

  f = open_out(dir, "out_ghost_%d_0.vtr", step);
  if(f == NULL) return -1;
  fprintf(f, "<VTKFile type=\"RectilinearGrid\">\n");
  fprintf(f, "<RectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\" "
    "GhostLevel=\"0\">\n", wx, wy, wz);
  fprintf(f, "<Piece Extent=\"%d %d %d %d %d %d\">\n", xs0, xs0 + xl, ys0,
    ys0 + yl, zs0, zs0 + zl);
  fprintf(f, "<CellData Scalars=\"p phase phase_shell\" "
    "Vectors=\"vel vel_star flag\">\n");
  /* the cell loops cover the piece; with one subdomain that is every cell
   * of the ghosted grid, in k, j, i order */
  const int k0 = zs0 + gc->ksb, k1 = zs0 + zl + gc->ksb;
  const int j0 = ys0 + gc->jsb, j1 = ys0 + yl + gc->jsb;
  const int i0 = xs0 + gc->isb, i1 = xs0 + xl + gc->isb;
  if(k1 > gc->knb || j1 > gc->jnb || i1 > gc->inb) {
    fprintf(stderr, "out_ghost: piece extent beyond the ghosted grid\n");
    fclose(f);
    return -1;
  }
  fprintf(f, "<DataArray type=\"Float32\" Name=\"p\">\n");
  for(int k = k0; k < k1; k++)
    for(int j = j0; j < j1; j++)
      for(int i = i0; i < i1; i++)
        fprintf(f, "%e ", p[i + j * gc->s1b + k * gc->s2b]);
  fprintf(f, "\n</DataArray>\n");
  const int *iv[2] = {phase, phase_shell};
  const char *inames[2] = {"phase", "phase_shell"};
  for(int q = 0; q < 2; q++) {
    fprintf(f, "<DataArray type=\"Int32\" Name=\"%s\">\n", inames[q]);
    for(int k = k0; k < k1; k++)
      for(int j = j0; j < j1; j++)
        for(int i = i0; i < i1; i++)
          fprintf(f, "%d ", iv[q][i + j * gc->s1b + k * gc->s2b]);
    fprintf(f, "\n</DataArray>\n");
  }
  fprintf(f, "<DataArray type=\"Float32\" Name=\"vel\" NumberOfComponents=\"3\""
    " format=\"ascii\">\n");
  const cell_range rg = {i0, i1, j0, j1, k0, k1};
  centre3(f, u, v, w, rg);
  fprintf(f, "\n</DataArray>\n");
  fprintf(f, "<DataArray type=\"Float32\" Name=\"vel_star\" "
    "NumberOfComponents=\"3\" format=\"ascii\">\n");
  centre3(f, u_star, v_star, w_star, rg);
  fprintf(f, "\n</DataArray>\n");
  fprintf(f, "<DataArray type=\"Float32\" Name=\"flag\" NumberOfComponents=\"3\""
    " format=\"ascii\">\n");
  centre3i(f, flag_u, flag_v, flag_w, rg);
  fprintf(f, "\n</DataArray>\n</CellData>\n<Coordinates>\n");
  const real s0[3] = {sd->xs, sd->ys, sd->zs};
  const real h[3] = {Dom.dx, Dom.dy, Dom.dz};
  const int n[3] = {xl, yl, zl};
  for(int a = 0; a < 3; a++) {
    fprintf(f, "<DataArray type=\"Float32\" Name=\"%c\">\n", "xyz"[a]);
    for(int q = -1; q < n[a]; q++) fprintf(f, "%lf ", s0[a] + h[a] * q);
    fprintf(f, "\n</DataArray>\n");
  }
  fprintf(f, "</Coordinates>\n</Piece>\n</RectilinearGrid>\n</VTKFile>\n");
  fclose(f);
  return 0;
}

int bb_vtk_ghost_init(const char *dir) { return pvd_init(dir, "out_ghost.pvd"); }

/* ---------------------------------------------------------------------
 * Particles: out_N.vtp
 * ------------------------------------------------------------------- */
static void vec_array(FILE *f, const char *name, const part_struct *pv, int n,
  void (*get)(const part_struct *, real *))
{
  fprintf(f, "<DataArray type=\"Float32\" NumberOfComponents=\"3\"  Name=\"%s\""
    " format=\"ascii\">\n", name);
  for(int q = 0; q < n; q++) {
    real v3[3];
    get(&pv[q], v3);
    fprintf(f, "%f %f %f ", v3[0], v3[1], v3[2]);
  }
  fprintf(f, "\n</DataArray>\n");
}

#define GETTER(nm, a, b, c) \
  static void nm(const part_struct *P, real *o) \
  { o[0] = (a); o[1] = (b); o[2] = (c); }
GETTER(get_vel, P->u, P->v, P->w)
GETTER(get_acc, P->udot, P->vdot, P->wdot)
GETTER(get_ax, P->axx, P->axy, P->axz)
GETTER(get_ay, P->ayx, P->ayy, P->ayz)
GETTER(get_az, P->azx, P->azy, P->azz)
GETTER(get_om, P->ox, P->oy, P->oz)
GETTER(get_omdot, P->oxdot, P->oydot, P->ozdot)
GETTER(get_L, P->Lx + P->iLx, P->Ly + P->iLy, P->Lz + P->iLz)
GETTER(get_Fh, P->Fx, P->Fy, P->Fz)
GETTER(get_Fk, P->kFx, P->kFy, P->kFz)
GETTER(get_Fi, P->iFx, P->iFy, P->iFz)

/* hydrodynamic + interaction + spring force + buoyant weight */
static void get_F(const part_struct *P, real *o)
{
  const real mass = 4. / 3. * VTK_PI * (P->rho - rho_f) * P->r * P->r * P->r;
  o[0] = P->Fx + P->iFx + P->kFx + mass * g.x;
  o[1] = P->Fy + P->iFy + P->kFy + mass * g.y;
  o[2] = P->Fz + P->iFz + P->kFz + mass * g.z;
}

static void piece_header(FILE *f, int npoints)
{
  fprintf(f, "<VTKFile type=\"PolyData\">\n<PolyData>\n");
  fprintf(f, "<Piece NumberOfPoints=\"%d\" NumberOfVerts=\"0\" "
    "NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n", npoints);
  fprintf(f, "<Points>\n<DataArray type=\"Float32\" NumberOfComponents=\"3\" "
    "format=\"ascii\">\n");
}

static const char *PIECE_END =
  "</PointData>\n</Piece>\n</PolyData>\n</VTKFile>\n";

int bb_vtk_parts(const char *dir, int step)
{
  part_struct *pv;
  int *ind;
  const int n = virtual_parts(&pv, &ind);
  if(n < 0) return -1;
  FILE *f = open_out(dir, "out_%d.vtp", step);
  if(f == NULL) { free(pv); free(ind); return -1; }
  piece_header(f, n);
  for(int q = 0; q < n; q++)
    fprintf(f, "%f %f %f ", pv[q].x, pv[q].y, pv[q].z);
  fprintf(f, "\n</DataArray>\n</Points>\n");
  fprintf(f, "<PointData Scalars=\"n r\" Vectors=\"pvel pacc pax pay paz "
    "pomega pomegadot F L Fh Fk Fi\">\n");
  fprintf(f, "<DataArray type=\"Int32\" Name=\"n\" format=\"ascii\">\n");
  for(int q = 0; q < n; q++) fprintf(f, "%d ", ind[q]);
  fprintf(f, "\n</DataArray>\n");
  fprintf(f, "<DataArray type=\"Float32\" Name=\"r\" format=\"ascii\">\n");
  for(int q = 0; q < n; q++) fprintf(f, "%f ", pv[q].r);
  fprintf(f, "\n</DataArray>\n");
  vec_array(f, "pvel", pv, n, get_vel);
  vec_array(f, "pacc", pv, n, get_acc);
  vec_array(f, "pax", pv, n, get_ax);
  vec_array(f, "pay", pv, n, get_ay);
  vec_array(f, "paz", pv, n, get_az);
  vec_array(f, "pomega", pv, n, get_om);
  vec_array(f, "pomegadot", pv, n, get_omdot);
  vec_array(f, "F", pv, n, get_F);
  vec_array(f, "L", pv, n, get_L);
  vec_array(f, "Fh", pv, n, get_Fh);
  vec_array(f, "Fk", pv, n, get_Fk);
  vec_array(f, "Fi", pv, n, get_Fi);
  fputs(PIECE_END, f);
  fclose(f);
  free(pv);
  free(ind);
  return 0;
}

/* ---------------------------------------------------------------------
 * Quadrature nodes: out_nodes_N.vtp
 * ------------------------------------------------------------------- */
/* the 26 node directions (theta, phi) of the degree-7 Lebedev rule in the
 * order of the quadrature: octahedral (poles nudged off the axis by
 * DIV_ST), edge, cube */
static void node_angles(real *t, real *ph)
{
  /* multiples of pi (each angle one rounding of c * pi); the poles and
   * the cube nodes' polar angles acos(+-1/sqrt(3)) are set apart */
  static const double TC[NNODES] = {0.5, 0.5, 0.5, 0.5, 0., 1.,
    0.5, 0.5, 0.5, 0.5, 0.25, 0.25, 0.25, 0.25, 0.75, 0.75, 0.75, 0.75,
    0., 0., 0., 0., 0., 0., 0., 0.};
  static const double PC[NNODES] = {0., 0.5, 1., 1.5, 0., 0.,
    0.25, 0.75, 1.25, 1.75, 0., 0.5, 1., 1.5, 0., 0.5, 1., 1.5,
    0.25, 0.75, 1.25, 1.75, 0.25, 0.75, 1.25, 1.75};
  for(int m = 0; m < NNODES; m++) {
    t[m] = TC[m] * VTK_PI;
    ph[m] = PC[m] * VTK_PI;
  }
  t[4] = 0. + DIV_ST;
  t[5] = VTK_PI - DIV_ST;
  for(int m = 18; m < 22; m++) t[m] = 0.955316618124509;
  for(int m = 22; m < 26; m++) t[m] = 2.186276035465284;
}

int bb_vtk_nodes(const char *dir, int step)
{
  part_struct *pv;
  int *ind;
  const int n = virtual_parts(&pv, &ind);
  if(n < 0) return -1;
  FILE *f = open_out(dir, "out_nodes_%d.vtp", step);
  if(f == NULL) { free(pv); free(ind); return -1; }
  real nt[NNODES], np[NNODES];
  node_angles(nt, np);
  const dom_struct *sd = &dom[0];
  piece_header(f, NNODES * n);
  for(int q = 0; q < n; q++) {
    const real rs = parts[ind[q]].rs;
    for(int m = 0; m < NNODES; m++) {
      real X = rs * sin(nt[m]) * cos(np[m]) + pv[q].x;
      real Y = rs * sin(nt[m]) * sin(np[m]) + pv[q].y;
      real Z = rs * cos(nt[m]) + pv[q].z;
      if(X < sd->xs) X += sd->xl;
      else if(X > sd->xe) X -= sd->xl;
      if(Y < sd->ys) Y += sd->yl;
      else if(Y > sd->ye) Y -= sd->yl;
      if(Z < sd->zs) Z += sd->zl;
      else if(Z > sd->xe) Z -= sd->zl;
      fprintf(f, "%f %f %f ", X, Y, Z);
    }
  }
  fprintf(f, "\n</DataArray>\n</Points>\n");
  fprintf(f, "<PointData Scalars=\"n status\">\n");
  fprintf(f, "<DataArray type=\"Int32\" Name=\"n\" format=\"ascii\">\n");
  for(int q = 0; q < n; q++)
    for(int m = 0; m < NNODES; m++) fprintf(f, "%d ", ind[q]);
  fprintf(f, "\n</DataArray>\n");
  fprintf(f, "<DataArray type=\"Int32\" Name=\"status\" format=\"ascii\">\n");
  for(int q = 0; q < n; q++)
    for(int m = 0; m < NNODES; m++) fprintf(f, "%d ", pv[q].nodes[m]);
  fprintf(f, "\n</DataArray>\n");
  fputs(PIECE_END, f);
  fclose(f);
  free(pv);
  free(ind);
  return 0;
}

/* one ParaView output: the collection entry and the three files */
int bb_vtk_out(const char *dir, int step)
{
  static const char *const kind[3] = {"out_%d.pvtr", "out_%d.vtp",
    "out_nodes_%d.vtp"};
  if(pvd_append(dir, "out.pvd", step, kind, 3)) return -1;
  if(bb_vtk_grid(dir, step)) return -1;
  if(bb_vtk_parts(dir, step)) return -1;
  return bb_vtk_nodes(dir, step);
}

/* one precursor-domain output: its collection entry and the grid files */
int bb_vtk_turb_out(const char *dir, int step)
{
  static const char *const kind[1] = {"out_turb_%d.pvtr"};
  if(pvd_append(dir, "out_turb.pvd", step, kind, 1)) return -1;
  return bb_vtk_turb_grid(dir, step);
}

/* one debugging output: the ghosted grid and the particles */
int bb_vtk_ghost_out(const char *dir, int step)
{
  static const char *const kind[2] = {"out_ghost_%d.pvtr", "out_%d.vtp"};
  if(pvd_append(dir, "out_ghost.pvd", step, kind, 2)) return -1;
  if(bb_vtk_ghost_grid(dir, step)) return -1;
  return bb_vtk_parts(dir, step);
}

/* ---- the reference's names, on ROOT_DIR/output, exiting on error ---- */
static void output_dir(char *buf, size_t len)
{
  snprintf(buf, len, "%s/output", bb_root_dir);
}

void init_VTK(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_init(d)) exit(EXIT_FAILURE);
}

void out_VTK(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_out(d, rec_paraview_stepnum_out)) exit(EXIT_FAILURE);
}

void dom_out_VTK(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_grid(d, rec_paraview_stepnum_out)) exit(EXIT_FAILURE);
}

void part_out_VTK(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_parts(d, rec_paraview_stepnum_out)) exit(EXIT_FAILURE);
}

void quadnodes_out_VTK(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_nodes(d, rec_paraview_stepnum_out)) exit(EXIT_FAILURE);
}

void init_VTK_turb(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_turb_init(d)) exit(EXIT_FAILURE);
}

void out_VTK_turb(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_turb_out(d, rec_prec_stepnum_out)) exit(EXIT_FAILURE);
}

void dom_out_VTK_turb(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_turb_grid(d, rec_prec_stepnum_out)) exit(EXIT_FAILURE);
}

void init_VTK_ghost(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_ghost_init(d)) exit(EXIT_FAILURE);
}

void out_VTK_ghost(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_ghost_out(d, rec_paraview_stepnum_out)) exit(EXIT_FAILURE);
}

void dom_out_VTK_ghost(void)
{
  char d[2 * FILE_NAME_SIZE];
  output_dir(d, sizeof(d));
  if(bb_vtk_ghost_grid(d, rec_paraview_stepnum_out)) exit(EXIT_FAILURE);
}
