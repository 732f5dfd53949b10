/*
 * rp_oracle.c -- CPU restatement of RePlAce's global-placement hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See rp_oracle.h for scope, the
 * list of reference files restated and how each part is pinned.  Numbered
 * comments cite reference (The-OpenROAD-Project/RePlAce) file:line.
 *
 * Floating-point rules: this file is compiled with -O2 -ffp-contract=off and
 * without -ffast-math so that every float / double expression rounds as the
 * reference's x86-64 build does (SSE2, no FMA).  long double is used where
 * the reference uses REPLACE_SQRT2 / REPLACE_FFT_PI (long double literals).
 */
#include "rp_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define RO_SQRT2 1.414213562373095048801L /* src/nesterovBase.cpp:13 */
#define RO_PI 3.141592653589793238462L    /* src/fft.cpp:9 */

#define NEW(T, n) ((T*)calloc((size_t)((n) > 0 ? (n) : 1), sizeof(T)))

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ */
/* MT19937 (Matsumoto & Nishimura 1998), the engine std::mt19937 fixes  */
/* ([rand.predef]); used by initFillerGCells, src/nesterovBase.cpp:1139 */
/* ------------------------------------------------------------------ */
typedef struct {
  uint32_t s[624];
  int i;
} mt_state;

static void mt_seed(mt_state* m, uint32_t seed) {
  m->s[0] = seed;
  for (int k = 1; k < 624; k++)
    m->s[k] = 1812433253u * (m->s[k - 1] ^ (m->s[k - 1] >> 30)) + (uint32_t)k;
  m->i = 624;
}

static uint32_t mt_next(mt_state* m) {
  if (m->i >= 624) {
    for (int k = 0; k < 624; k++) {
      uint32_t y = (m->s[k] & 0x80000000u) | (m->s[(k + 1) % 624] & 0x7fffffffu);
      uint32_t v = m->s[(k + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      m->s[k] = v;
    }
    m->i = 0;
  }
  uint32_t y = m->s[m->i++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

void ro_mt19937(uint32_t seed, uint32_t* out, int n) {
  mt_state m;
  mt_seed(&m, seed);
  for (int k = 0; k < n; k++) out[k] = mt_next(&m);
}

/* fastExp, src/nesterovBase.cpp:1568-1582: (1 + a/1024) in double rounded
 * to float, then ten float squarings. */
float ro_fast_exp(float a) {
  a = (float)(1.0 + (double)a / 1024.0);
  for (int k = 0; k < 10; k++) a = a * a;
  return a;
}

/* ------------------------------------------------------------------ */
/* Context                                                            */
/* ------------------------------------------------------------------ */
struct ro_ctx {
  rp_params p;
  /* wall time per stage of the Nesterov loop (instrumentation only) */
  double t_stage[RO_STAGES];
  /* die / core, src/placerBase.cpp:614-626 */
  int core_lx, core_ly, core_ux, core_uy;
  int site_w, site_h;
  /* instances: real ones first, then fragmented-row dummies */
  int n_real, n_inst;
  int *ilx, *ily, *iux, *iuy; /* Instance box (src/placerBase.cpp:41-55) */
  int *mw, *mh;               /* master size (DB bbox) */
  int *dblx, *dbly;           /* DB location */
  uint8_t* ifixed;
  int* place;
  int n_place;
  int* inst2place;
  int* nonplace;
  int n_nonplace;
  int64_t place_area, nonplace_area, macro_area, std_area;
  /* pins / nets */
  int n_pin, n_net;
  int* net_start;
  int* pin_inst;
  int *pin_offx, *pin_offy;
  int *pin_cx, *pin_cy;
  int* pin_bbox; /* 4 per pin for BTerms */
  int *ipin_start, *ipins;
  uint8_t* pin_mm; /* bit0 minX bit1 maxX bit2 minY bit3 maxY */
  /* initial place */
  float *ipx, *ipy;
  /* Nesterov base */
  int nb_ready;
  int n_g, n_gi, n_gf;
  int* g_inst;
  uint8_t* g_macro;
  int *g_lx, *g_ly, *g_ux, *g_uy;
  int *g_dlx, *g_dly, *g_dux, *g_duy;
  float* g_scale;
  int *g_pstart, *g_pins;
  int* gp_cell;
  int *gp_offx, *gp_offy;
  int *gp_cx, *gp_cy;
  float* gp_exp;     /* 4 per pin: minX maxX minY maxY */
  uint8_t* gp_has;   /* bit0 minX bit1 maxX bit2 minY bit3 maxY */
  int* gn_box;       /* 4 per net */
  float* gn_sum;     /* 8 per net */
  int filler_dx, filler_dy;
  /* bin grid */
  int nbx, nby, bsx, bsy;
  float td;
  int *b_lx, *b_ly, *b_ux, *b_uy;
  int64_t *b_nonplace, *b_inst, *b_filler;
  float *b_density, *b_phi, *b_ex, *b_ey;
  int64_t overflow_area;
  float sum_phi;
  /* Poisson */
  ro_poisson_fn pfn;
  void* puser;
  float *fft_wx, *fft_wy; /* unused by hooks; kept for the own transform */
  /* Nesterov place */
  int np_ready;
  float *curSLP, *curSLPWL, *curSLPD, *curSLPSum;
  float *nextSLP, *nextSLPWL, *nextSLPD, *nextSLPSum;
  float *prevSLP, *prevSLPWL, *prevSLPD, *prevSLPSum;
  float *curCo, *nextCo;
  float wlGradSum, densGradSum, stepLength, densityPenalty;
  float baseWLCoef, wlCoefX, wlCoefY, sumOverflow;
  int64_t prevHpwl;
  int isDiverged;
  float maxPhiCoef;
  /* loop state (doNesterovPlace locals, kept across ro_nesterov_iterate) */
  float curA, minSumOverflow, hpwlWithMinSumOverflow;
  int isMaxPhiCoefChanged, iter, finished, code;
  rp_iter_record* rec;
  int n_rec;
  int evals;
};

/* ------------------------------------------------------------------ */
/* PlacerBase (src/placerBase.cpp)                                    */
/* ------------------------------------------------------------------ */
static void pair_minmax_idx(int ll, int uu, int coreLL, int size, int minIdx, int maxIdx,
                            int* lo, int* hi) {
  /* getMinMaxIdx, src/placerBase.cpp:984-993 */
  int lower = (ll - coreLL) / size;
  int rem = uu - coreLL;
  int fm = rem >= size ? rem % size : rem;
  int upper = (fm == 0) ? rem / size : rem / size + 1;
  *lo = imax(minIdx, lower);
  *hi = imin(maxIdx, upper);
}

static int inst_cx(ro_ctx* c, int i) { return (c->ilx[i] + c->iux[i]) / 2; }
static int inst_cy(ro_ctx* c, int i) { return (c->ily[i] + c->iuy[i]) / 2; }

static void inst_update_pins(ro_ctx* c, int i) {
  /* Pin::updateLocation, src/placerBase.cpp:408-412 */
  for (int k = c->ipin_start[i]; k < c->ipin_start[i + 1]; k++) {
    int p = c->ipins[k];
    c->pin_cx[p] = inst_cx(c, i) + c->pin_offx[p];
    c->pin_cy[p] = inst_cy(c, i) + c->pin_offy[p];
  }
}

static void inst_set_center(ro_ctx* c, int i, int x, int y) {
  /* Instance::setCenterLocation, src/placerBase.cpp:127-140 */
  int hx = (c->iux[i] - c->ilx[i]) / 2;
  int hy = (c->iuy[i] - c->ily[i]) / 2;
  c->ilx[i] = x - hx;
  c->ily[i] = y - hy;
  c->iux[i] = x + hx;
  c->iuy[i] = y + hy;
  inst_update_pins(c, i);
}

ro_ctx* ro_create(const rp_netlist* nl, const rp_params* p, int* code) {
  *code = 0;
  ro_ctx* c = NEW(ro_ctx, 1);
  c->p = *p;
  /* core = bbox of rows, src/placerBase.cpp:961-976 */
  int minX = INT_MAX, minY = INT_MAX, maxX = INT_MIN, maxY = INT_MIN;
  for (int r = 0; r < nl->num_rows; r++) {
    minX = imin(nl->row_box[4 * r + 0], minX);
    minY = imin(nl->row_box[4 * r + 1], minY);
    maxX = imax(nl->row_box[4 * r + 2], maxX);
    maxY = imax(nl->row_box[4 * r + 3], maxY);
  }
  c->core_lx = minX;
  c->core_ly = minY;
  c->core_ux = maxX;
  c->core_uy = maxY;
  c->site_w = nl->site_w;
  c->site_h = nl->site_h;

  int nr = nl->num_insts;
  /* fragmented-row dummies, src/placerBase.cpp:760-846 */
  int scx = (c->core_ux - c->core_lx) / c->site_w;
  int scy = (c->core_uy - c->core_ly) / c->site_h;
  uint8_t* grid = (uint8_t*)calloc((size_t)scx * (size_t)scy + 1, 1); /* 0 Empty 1 Row 2 Fixed */
  for (int r = 0; r < nl->num_rows; r++) {
    int x0, x1, y0, y1;
    pair_minmax_idx(nl->row_box[4 * r + 0], nl->row_box[4 * r + 2], c->core_lx, c->site_w, 0, scx,
                    &x0, &x1);
    pair_minmax_idx(nl->row_box[4 * r + 1], nl->row_box[4 * r + 3], c->core_ly, c->site_h, 0, scy,
                    &y0, &y1);
    for (int i = x0; i < x1; i++)
      for (int j = y0; j < y1; j++) grid[(size_t)j * scx + i] = 1;
  }
  for (int i = 0; i < nr; i++) {
    if (!nl->inst_fixed[i]) continue;
    int x0, x1, y0, y1;
    pair_minmax_idx(nl->inst_box[4 * i + 0], nl->inst_box[4 * i + 2], c->core_lx, c->site_w, 0,
                    scx, &x0, &x1);
    pair_minmax_idx(nl->inst_box[4 * i + 1], nl->inst_box[4 * i + 3], c->core_ly, c->site_h, 0,
                    scy, &y0, &y1);
    for (int a = x0; a < x1; a++)
      for (int b = y0; b < y1; b++) grid[(size_t)b * scx + a] = 2;
  }
  int n_dummy = 0, cap_dummy = 16;
  int* dbox = NEW(int, 4 * cap_dummy);
  for (int j = 0; j < scy; j++) {
    for (int i = 0; i < scx; i++) {
      if (grid[(size_t)j * scx + i] == 0) {
        int sx = i;
        while (i < scx && grid[(size_t)j * scx + i] == 0) i++;
        int ex = i;
        if (n_dummy == cap_dummy) {
          cap_dummy *= 2;
          dbox = (int*)realloc(dbox, sizeof(int) * 4 * cap_dummy);
        }
        dbox[4 * n_dummy + 0] = c->core_lx + c->site_w * sx;
        dbox[4 * n_dummy + 1] = c->core_ly + c->site_h * j;
        dbox[4 * n_dummy + 2] = c->core_lx + c->site_w * ex;
        dbox[4 * n_dummy + 3] = c->core_ly + c->site_h * (j + 1);
        n_dummy++;
      }
    }
  }
  free(grid);

  c->n_real = nr;
  c->n_inst = nr + n_dummy;
  int ni = c->n_inst;
  c->ilx = NEW(int, ni);
  c->ily = NEW(int, ni);
  c->iux = NEW(int, ni);
  c->iuy = NEW(int, ni);
  c->mw = NEW(int, ni);
  c->mh = NEW(int, ni);
  c->dblx = NEW(int, ni);
  c->dbly = NEW(int, ni);
  c->ifixed = NEW(uint8_t, ni);
  c->inst2place = NEW(int, ni);
  c->place = NEW(int, ni);
  c->nonplace = NEW(int, ni);
  for (int i = 0; i < nr; i++) {
    c->ilx[i] = nl->inst_box[4 * i + 0];
    c->ily[i] = nl->inst_box[4 * i + 1];
    c->iux[i] = nl->inst_box[4 * i + 2];
    c->iuy[i] = nl->inst_box[4 * i + 3];
    c->mw[i] = c->iux[i] - c->ilx[i];
    c->mh[i] = c->iuy[i] - c->ily[i];
    c->dblx[i] = c->ilx[i];
    c->dbly[i] = c->ily[i];
    c->ifixed[i] = nl->inst_fixed[i] ? 1 : 0;
  }
  for (int d = 0; d < n_dummy; d++) {
    int i = nr + d;
    c->ilx[i] = dbox[4 * d + 0];
    c->ily[i] = dbox[4 * d + 1];
    c->iux[i] = dbox[4 * d + 2];
    c->iuy[i] = dbox[4 * d + 3];
    c->ifixed[i] = 1;
  }
  free(dbox);
  /* classification and areas, src/placerBase.cpp:639-680 */
  for (int i = 0; i < ni; i++) {
    c->inst2place[i] = -1;
    if (i < nr) {
      if (c->ifixed[i]) {
        int rlx = imax(c->core_lx, c->ilx[i]), rly = imax(c->core_ly, c->ily[i]);
        int rux = imin(c->core_ux, c->iux[i]), ruy = imin(c->core_uy, c->iuy[i]);
        if (!(rlx >= rux || rly >= ruy)) {
          c->nonplace[c->n_nonplace++] = i;
          c->nonplace_area += (int64_t)(rux - rlx) * (int64_t)(ruy - rly);
        }
      } else {
        c->inst2place[i] = c->n_place;
        c->place[c->n_place++] = i;
        int64_t a = (int64_t)(c->iux[i] - c->ilx[i]) * (int64_t)(c->iuy[i] - c->ily[i]);
        c->place_area += a;
        if (c->iuy[i] - c->ily[i] > c->site_h * 6)
          c->macro_area += a;
        else
          c->std_area += a;
      }
    } else {
      c->nonplace[c->n_nonplace++] = i;
      c->nonplace_area += (int64_t)(c->iux[i] - c->ilx[i]) * (int64_t)(c->iuy[i] - c->ily[i]);
    }
  }
  /* pins, src/placerBase.cpp:702-713, 330-406 */
  c->n_net = nl->num_nets;
  c->n_pin = nl->num_pins;
  int np = c->n_pin;
  c->net_start = NEW(int, c->n_net + 1);
  memcpy(c->net_start, nl->net_pin_start, sizeof(int) * (c->n_net + 1));
  c->pin_inst = NEW(int, np);
  c->pin_offx = NEW(int, np);
  c->pin_offy = NEW(int, np);
  c->pin_cx = NEW(int, np);
  c->pin_cy = NEW(int, np);
  c->pin_bbox = NEW(int, 4 * np);
  c->pin_mm = NEW(uint8_t, np);
  for (int q = 0; q < np; q++) {
    int i = nl->pin_inst[q];
    c->pin_inst[q] = i;
    if (i >= 0) {
      c->pin_offx[q] = nl->pin_off[2 * q];
      c->pin_offy[q] = nl->pin_off[2 * q + 1];
      /* cx = inst lx + masterW/2 + offset, src/placerBase.cpp:345-375 */
      c->pin_cx[q] = c->ilx[i] + c->mw[i] / 2 + c->pin_offx[q];
      c->pin_cy[q] = c->ily[i] + c->mh[i] / 2 + c->pin_offy[q];
    } else {
      int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
      if (nl->pin_box) {
        lx = nl->pin_box[4 * q + 0];
        ly = nl->pin_box[4 * q + 1];
        ux = nl->pin_box[4 * q + 2];
        uy = nl->pin_box[4 * q + 3];
      }
      c->pin_bbox[4 * q + 0] = lx;
      c->pin_bbox[4 * q + 1] = ly;
      c->pin_bbox[4 * q + 2] = ux;
      c->pin_bbox[4 * q + 3] = uy;
      c->pin_offx[q] = c->pin_offy[q] = 0;
      /* (lx+ux)/2 in int; for an unplaced port INT_MAX+INT_MIN = -1 -> 0 */
      c->pin_cx[q] = (int)(((int64_t)lx + (int64_t)ux) / 2);
      c->pin_cy[q] = (int)(((int64_t)ly + (int64_t)uy) / 2);
    }
  }
  /* per-instance pin lists, src/placerBase.cpp:728-743 */
  c->ipin_start = NEW(int, ni + 1);
  c->ipins = NEW(int, np);
  if (nl->inst_pin_start && nl->inst_pins) {
    for (int i = 0; i <= nr; i++) c->ipin_start[i] = nl->inst_pin_start[i];
    for (int i = nr + 1; i <= ni; i++) c->ipin_start[i] = c->ipin_start[nr];
    memcpy(c->ipins, nl->inst_pins, sizeof(int) * nl->inst_pin_start[nr]);
  } else {
    int* cnt = NEW(int, ni + 1);
    for (int q = 0; q < np; q++)
      if (c->pin_inst[q] >= 0) cnt[c->pin_inst[q] + 1]++;
    for (int i = 0; i < ni; i++) cnt[i + 1] += cnt[i];
    memcpy(c->ipin_start, cnt, sizeof(int) * (ni + 1));
    for (int q = 0; q < np; q++)
      if (c->pin_inst[q] >= 0) c->ipins[cnt[c->pin_inst[q]]++] = q;
    free(cnt);
  }
  /* printInfo utilisation check, src/placerBase.cpp:936-956 */
  int64_t coreArea = (int64_t)(c->core_ux - c->core_lx) * (int64_t)(c->core_uy - c->core_ly);
  float util = (float)c->place_area / (float)(coreArea - c->nonplace_area) * 100;
  if (util >= 100.1) *code = 1;
  c->pfn = ro_poisson_fft64;
  c->puser = NULL;
  return c;
}

void ro_set_poisson(ro_ctx* c, ro_poisson_fn fn, void* user) {
  c->pfn = fn ? fn : ro_poisson_fft64;
  c->puser = user;
}

int ro_num_insts(ro_ctx* c) { return c->n_real; }

#define FREE(x) free(c->x)
void ro_destroy(ro_ctx* c) {
  if (!c) return;
  FREE(ilx); FREE(ily); FREE(iux); FREE(iuy); FREE(mw); FREE(mh); FREE(dblx); FREE(dbly);
  FREE(ifixed); FREE(place); FREE(inst2place); FREE(nonplace);
  FREE(net_start); FREE(pin_inst); FREE(pin_offx); FREE(pin_offy); FREE(pin_cx); FREE(pin_cy);
  FREE(pin_bbox); FREE(ipin_start); FREE(ipins); FREE(pin_mm); FREE(ipx); FREE(ipy);
  FREE(g_inst); FREE(g_macro); FREE(g_lx); FREE(g_ly); FREE(g_ux); FREE(g_uy);
  FREE(g_dlx); FREE(g_dly); FREE(g_dux); FREE(g_duy); FREE(g_scale); FREE(g_pstart); FREE(g_pins);
  FREE(gp_cell); FREE(gp_offx); FREE(gp_offy); FREE(gp_cx); FREE(gp_cy); FREE(gp_exp); FREE(gp_has);
  FREE(gn_box); FREE(gn_sum); FREE(b_lx); FREE(b_ly); FREE(b_ux); FREE(b_uy);
  FREE(b_nonplace); FREE(b_inst); FREE(b_filler); FREE(b_density); FREE(b_phi); FREE(b_ex);
  FREE(b_ey); FREE(fft_wx); FREE(fft_wy);
  FREE(curSLP); FREE(curSLPWL); FREE(curSLPD); FREE(curSLPSum);
  FREE(nextSLP); FREE(nextSLPWL); FREE(nextSLPD); FREE(nextSLPSum);
  FREE(prevSLP); FREE(prevSLPWL); FREE(prevSLPD); FREE(prevSLPSum);
  FREE(curCo); FREE(nextCo); FREE(rec);
  free(c);
}

/* PlacerBase::hpwl over DB instance boxes and BPin boxes,
 * src/placerBase.cpp:868-876, Net::updateBox :480-500 */
int64_t ro_inst_hpwl(ro_ctx* c) {
  int64_t h = 0;
  for (int n = 0; n < c->n_net; n++) {
    int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
    for (int q = c->net_start[n]; q < c->net_start[n + 1]; q++) {
      int i = c->pin_inst[q];
      if (i < 0) continue;
      lx = imin(c->dblx[i], lx);
      ly = imin(c->dbly[i], ly);
      ux = imax(c->dblx[i] + c->mw[i], ux);
      uy = imax(c->dbly[i] + c->mh[i], uy);
    }
    for (int q = c->net_start[n]; q < c->net_start[n + 1]; q++) {
      if (c->pin_inst[q] >= 0) continue;
      if (c->pin_bbox[4 * q] == INT_MAX) continue; /* no BPin shapes */
      lx = imin(c->pin_bbox[4 * q + 0], lx);
      ly = imin(c->pin_bbox[4 * q + 1], ly);
      ux = imax(c->pin_bbox[4 * q + 2], ux);
      uy = imax(c->pin_bbox[4 * q + 3], uy);
    }
    h += (int64_t)((int)((unsigned)ux - (unsigned)lx) + (int)((unsigned)uy - (unsigned)ly));
  }
  return h;
}

/* ------------------------------------------------------------------ */
/* InitialPlace (src/initialPlace.cpp)                                */
/* ------------------------------------------------------------------ */
typedef struct {
  int n, nnz;
  int* rowp;
  int* col;
  float* val;
} csr_f;

typedef struct {
  int r, c;
  float v;
} trip;

static int trip_cmp(const void* a, const void* b) {
  const trip* x = (const trip*)a;
  const trip* y = (const trip*)b;
  if (x->r != y->r) return x->r < y->r ? -1 : 1;
  if (x->c != y->c) return x->c < y->c ? -1 : 1;
  return 0;
}

/* setFromTriplets: duplicates summed in insertion order (stable sort). */
static void csr_from_trips(csr_f* m, int n, trip* t, int nt, int* order_tmp) {
  (void)order_tmp;
  /* stable: sort with index tie-break */
  typedef struct {
    trip t;
    int k;
  } tk;
  tk* a = (tk*)malloc(sizeof(tk) * (size_t)(nt > 0 ? nt : 1));
  for (int k = 0; k < nt; k++) {
    a[k].t = t[k];
    a[k].k = k;
  }
  qsort(a, (size_t)nt, sizeof(tk), (int (*)(const void*, const void*))trip_cmp);
  /* qsort is not stable: re-sort equal (r,c) runs by k */
  for (int s = 0; s < nt;) {
    int e = s + 1;
    while (e < nt && a[e].t.r == a[s].t.r && a[e].t.c == a[s].t.c) e++;
    for (int x = s + 1; x < e; x++) {
      tk v = a[x];
      int y = x - 1;
      while (y >= s && a[y].k > v.k) {
        a[y + 1] = a[y];
        y--;
      }
      a[y + 1] = v;
    }
    s = e;
  }
  m->n = n;
  m->rowp = NEW(int, n + 1);
  m->col = NEW(int, nt);
  m->val = NEW(float, nt);
  int nnz = 0;
  for (int s = 0; s < nt;) {
    int e = s;
    float v = 0.0f;
    while (e < nt && a[e].t.r == a[s].t.r && a[e].t.c == a[s].t.c) {
      v += a[e].t.v;
      e++;
    }
    m->col[nnz] = a[s].t.c;
    m->val[nnz] = v;
    m->rowp[a[s].t.r + 1]++;
    nnz++;
    s = e;
  }
  for (int r = 0; r < n; r++) m->rowp[r + 1] += m->rowp[r];
  m->nnz = nnz;
  free(a);
}

static void csr_free(csr_f* m) {
  free(m->rowp);
  free(m->col);
  free(m->val);
}

static void spmv(const csr_f* m, const float* x, float* y) {
  for (int r = 0; r < m->n; r++) {
    double s = 0.0;
    for (int k = m->rowp[r]; k < m->rowp[r + 1]; k++) s += (double)m->val[k] * (double)x[m->col[k]];
    y[r] = (float)s;
  }
}

static double dotd(const float* a, const float* b, int n) {
  double s = 0.0;
  for (int k = 0; k < n; k++) s += (double)a[k] * (double)b[k];
  return s;
}

/* BiCGSTAB with identity preconditioner, the published algorithm
 * (van der Vorst 1992) in the variant with a restart when rho collapses.
 * float vectors, FP64 dot products.  Returns the relative residual.
 * Reference call site: src/initialPlace.cpp:78-86 (Eigen, absent here). */
static float bicgstab(const csr_f* A, const float* b, float* x, int max_iter, float tol) {
  int n = A->n;
  float* r = NEW(float, n);
  float* r0 = NEW(float, n);
  float* v = NEW(float, n);
  float* pv = NEW(float, n);
  float* s = NEW(float, n);
  float* t = NEW(float, n);
  spmv(A, x, t);
  for (int k = 0; k < n; k++) r[k] = b[k] - t[k];
  memcpy(r0, r, sizeof(float) * n);
  double r0_sq = dotd(r0, r0, n);
  double rhs_sq = dotd(b, b, n);
  float err;
  if (rhs_sq == 0.0) {
    for (int k = 0; k < n; k++) x[k] = 0.0f;
    err = 0.0f;
    goto done;
  }
  {
    double rho = 1.0, alpha = 1.0, w = 1.0;
    double tol2 = (double)tol * (double)tol * rhs_sq;
    double eps = 1.1920928955078125e-07;
    double eps2 = eps * eps;
    int i = 0, restarts = 0;
    while (dotd(r, r, n) > tol2 && i < max_iter) {
      double rho_old = rho;
      rho = dotd(r0, r, n);
      if (fabs(rho) < eps2 * r0_sq) {
        spmv(A, x, t);
        for (int k = 0; k < n; k++) r[k] = b[k] - t[k];
        memcpy(r0, r, sizeof(float) * n);
        rho = r0_sq = dotd(r, r, n);
        if (restarts++ == 0) i = 0;
      }
      double beta = (rho / rho_old) * (alpha / w);
      for (int k = 0; k < n; k++) pv[k] = (float)((double)r[k] + beta * ((double)pv[k] - w * (double)v[k]));
      spmv(A, pv, v);
      alpha = rho / dotd(r0, v, n);
      for (int k = 0; k < n; k++) s[k] = (float)((double)r[k] - alpha * (double)v[k]);
      spmv(A, s, t);
      double tt = dotd(t, t, n);
      w = tt > 0.0 ? dotd(t, s, n) / tt : 0.0;
      for (int k = 0; k < n; k++) {
        x[k] = (float)((double)x[k] + alpha * (double)pv[k] + w * (double)s[k]);
        r[k] = (float)((double)s[k] - w * (double)t[k]);
      }
      i++;
    }
    err = (float)sqrt(dotd(r, r, n) / rhs_sq);
  }
done:
  free(r); free(r0); free(v); free(pv); free(s); free(t);
  return err;
}

int ro_initial_place(ro_ctx* c, float* err_out, int64_t* hpwl_out, int cap, int* n_iter) {
  const rp_params* p = &c->p;
  int n = c->n_place;
  *n_iter = 0;
  /* placeInstsCenter, src/initialPlace.cpp:108-115 */
  if (!p->incremental_place_mode) {
    int ccx = (c->core_lx + c->core_ux) / 2, ccy = (c->core_ly + c->core_uy) / 2;
    for (int k = 0; k < n; k++) inst_set_center(c, c->place[k], ccx, ccy);
  }
  free(c->ipx);
  free(c->ipy);
  c->ipx = NEW(float, n);
  c->ipy = NEW(float, n);
  float* bx = NEW(float, n);
  float* by = NEW(float, n);
  int cap_t = 1024;
  trip* tx = (trip*)malloc(sizeof(trip) * cap_t);
  trip* ty = (trip*)malloc(sizeof(trip) * cap_t);
  for (int it = 1; it <= p->initial_place_max_iter; it++) {
    /* updatePinInfo, src/initialPlace.cpp:128-182 */
    memset(c->pin_mm, 0, (size_t)c->n_pin);
    for (int e = 0; e < c->n_net; e++) {
      int pminx = -1, pmaxx = -1, pminy = -1, pmaxy = -1;
      int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
      for (int q = c->net_start[e]; q < c->net_start[e + 1]; q++) {
        if (lx > c->pin_cx[q]) {
          if (pminx >= 0) c->pin_mm[pminx] &= (uint8_t)~1u;
          lx = c->pin_cx[q];
          pminx = q;
          c->pin_mm[q] |= 1u;
        }
        if (ux < c->pin_cx[q]) {
          if (pmaxx >= 0) c->pin_mm[pmaxx] &= (uint8_t)~2u;
          ux = c->pin_cx[q];
          pmaxx = q;
          c->pin_mm[q] |= 2u;
        }
        if (ly > c->pin_cy[q]) {
          if (pminy >= 0) c->pin_mm[pminy] &= (uint8_t)~4u;
          ly = c->pin_cy[q];
          pminy = q;
          c->pin_mm[q] |= 4u;
        }
        if (uy < c->pin_cy[q]) {
          if (pmaxy >= 0) c->pin_mm[pmaxy] &= (uint8_t)~8u;
          uy = c->pin_cy[q];
          pmaxy = q;
          c->pin_mm[q] |= 8u;
        }
      }
    }
    /* createSparseMatrix, src/initialPlace.cpp:185-378 */
    for (int k = 0; k < n; k++) {
      int i = c->place[k];
      c->ipx[k] = (float)inst_cx(c, i);
      c->ipy[k] = (float)inst_cy(c, i);
      bx[k] = by[k] = 0.0f;
    }
    int ntx = 0, nty = 0;
    for (int e = 0; e < c->n_net; e++) {
      int s = c->net_start[e], deg = c->net_start[e + 1] - s;
      if (deg <= 1) continue;
      if (deg >= p->initial_place_max_fanout) continue;
      float netWeight = p->initial_place_net_weight_scale / (float)(deg - 1);
      for (int a1 = 0; a1 < deg; a1++) {
        int q1 = s + a1;
        for (int a2 = 0; a2 < deg; a2++) {
          int q2 = s + a2;
          if (a1 < a2) break;
          int i1 = c->pin_inst[q1], i2 = c->pin_inst[q2];
          if (i1 == i2) continue;
          int pc1 = (i1 >= 0) && c->inst2place[i1] >= 0;
          int pc2 = (i2 >= 0) && c->inst2place[i2] >= 0;
          if (ntx + 4 > cap_t || nty + 4 > cap_t) {
            cap_t *= 2;
            tx = (trip*)realloc(tx, sizeof(trip) * cap_t);
            ty = (trip*)realloc(ty, sizeof(trip) * cap_t);
          }
          if ((c->pin_mm[q1] & 3u) || (c->pin_mm[q2] & 3u)) {
            int diff = abs(c->pin_cx[q1] - c->pin_cx[q2]);
            float w = diff > p->initial_place_min_diff_length
                          ? netWeight / (float)diff
                          : netWeight / (float)p->initial_place_min_diff_length;
            if (pc1 && pc2) {
              int k1 = c->inst2place[i1], k2 = c->inst2place[i2];
              tx[ntx++] = (trip){k1, k1, w};
              tx[ntx++] = (trip){k2, k2, w};
              tx[ntx++] = (trip){k1, k2, -w};
              tx[ntx++] = (trip){k2, k1, -w};
              bx[k1] += -w * (float)((c->pin_cx[q1] - inst_cx(c, i1)) - (c->pin_cx[q2] - inst_cx(c, i2)));
              bx[k2] += -w * (float)((c->pin_cx[q2] - inst_cx(c, i2)) - (c->pin_cx[q1] - inst_cx(c, i1)));
            } else if (!pc1 && pc2) {
              int k2 = c->inst2place[i2];
              tx[ntx++] = (trip){k2, k2, w};
              bx[k2] += w * (float)(c->pin_cx[q1] - (c->pin_cx[q2] - inst_cx(c, i2)));
            } else if (pc1 && !pc2) {
              int k1 = c->inst2place[i1];
              tx[ntx++] = (trip){k1, k1, w};
              bx[k1] += w * (float)(c->pin_cx[q2] - (c->pin_cx[q1] - inst_cx(c, i1)));
            }
          }
          if ((c->pin_mm[q1] & 12u) || (c->pin_mm[q2] & 12u)) {
            int diff = abs(c->pin_cy[q1] - c->pin_cy[q2]);
            float w = diff > p->initial_place_min_diff_length
                          ? netWeight / (float)diff
                          : netWeight / (float)p->initial_place_min_diff_length;
            if (pc1 && pc2) {
              int k1 = c->inst2place[i1], k2 = c->inst2place[i2];
              ty[nty++] = (trip){k1, k1, w};
              ty[nty++] = (trip){k2, k2, w};
              ty[nty++] = (trip){k1, k2, -w};
              ty[nty++] = (trip){k2, k1, -w};
              by[k1] += -w * (float)((c->pin_cy[q1] - inst_cy(c, i1)) - (c->pin_cy[q2] - inst_cy(c, i2)));
              by[k2] += -w * (float)((c->pin_cy[q2] - inst_cy(c, i2)) - (c->pin_cy[q1] - inst_cy(c, i1)));
            } else if (!pc1 && pc2) {
              int k2 = c->inst2place[i2];
              ty[nty++] = (trip){k2, k2, w};
              by[k2] += w * (float)(c->pin_cy[q1] - (c->pin_cy[q2] - inst_cy(c, i2)));
            } else if (pc1 && !pc2) {
              int k1 = c->inst2place[i1];
              ty[nty++] = (trip){k1, k1, w};
              by[k1] += w * (float)(c->pin_cy[q2] - (c->pin_cy[q1] - inst_cy(c, i1)));
            }
          }
        }
      }
    }
    csr_f Ax, Ay;
    csr_from_trips(&Ax, n, tx, ntx, NULL);
    csr_from_trips(&Ay, n, ty, nty, NULL);
    float tol = 1.1920928955078125e-07f;
    float ex = bicgstab(&Ax, bx, c->ipx, p->initial_place_max_solver_iter, tol);
    float ey = bicgstab(&Ay, by, c->ipy, p->initial_place_max_solver_iter, tol);
    csr_free(&Ax);
    csr_free(&Ay);
    float err = ex > ey ? ex : ey;
    int64_t h = ro_inst_hpwl(c); /* printed before updateCoordi, :88-91 */
    if (it - 1 < cap) {
      err_out[it - 1] = err;
      hpwl_out[it - 1] = h;
    }
    *n_iter = it;
    /* updateCoordi, src/initialPlace.cpp:380-385 */
    for (int k = 0; k < n; k++) {
      int i = c->place[k];
      inst_set_center(c, i, (int)c->ipx[k], (int)c->ipy[k]);
      c->dblx[i] = c->ilx[i];
      c->dbly[i] = c->ily[i];
    }
    if (err <= 1e-5 && it >= 5) break;
  }
  free(bx);
  free(by);
  free(tx);
  free(ty);
  return 0;
}

/* ------------------------------------------------------------------ */
/* NesterovBase (src/nesterovBase.cpp)                                */
/* ------------------------------------------------------------------ */
static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static void gcell_set_density_size(ro_ctx* c, int g, int ddx, int ddy) {
  /* GCell::setDensitySize, src/nesterovBase.cpp:182-191 */
  int cx = (c->g_dux[g] + c->g_dlx[g]) / 2;
  int cy = (c->g_duy[g] + c->g_dly[g]) / 2;
  c->g_dlx[g] = cx - ddx / 2;
  c->g_dly[g] = cy - ddy / 2;
  c->g_dux[g] = cx + ddx / 2;
  c->g_duy[g] = cy + ddy / 2;
}

static void gcell_update_pins(ro_ctx* c, int g) {
  /* GPin::updateDensityLocation, src/nesterovBase.cpp:411-415 */
  int dcx = (c->g_dux[g] + c->g_dlx[g]) / 2;
  int dcy = (c->g_duy[g] + c->g_dly[g]) / 2;
  for (int k = c->g_pstart[g]; k < c->g_pstart[g + 1]; k++) {
    int q = c->g_pins[k];
    c->gp_cx[q] = dcx + c->gp_offx[q];
    c->gp_cy[q] = dcy + c->gp_offy[q];
  }
}

static void gcell_set_density_center(ro_ctx* c, int g, int x, int y) {
  /* GCell::setDensityCenterLocation, src/nesterovBase.cpp:164-179 */
  int hx = (c->g_dux[g] - c->g_dlx[g]) / 2;
  int hy = (c->g_duy[g] - c->g_dly[g]) / 2;
  c->g_dlx[g] = x - hx;
  c->g_dly[g] = y - hy;
  c->g_dux[g] = x + hx;
  c->g_duy[g] = y + hy;
  gcell_update_pins(c, g);
}

static void bins_nonplace(ro_ctx* c) {
  /* BinGrid::updateBinsNonPlaceArea, src/nesterovBase.cpp:719-741 */
  for (int b = 0; b < c->nbx * c->nby; b++) c->b_nonplace[b] = 0;
  for (int k = 0; k < c->n_nonplace; k++) {
    int i = c->nonplace[k];
    int x0, x1, y0, y1;
    pair_minmax_idx(c->ilx[i], c->iux[i], c->core_lx, c->bsx, 0, c->nbx, &x0, &x1);
    pair_minmax_idx(c->ily[i], c->iuy[i], c->core_ly, c->bsy, 0, c->nby, &y0, &y1);
    for (int a = x0; a < x1; a++) {
      for (int b = y0; b < y1; b++) {
        int idx = b * c->nbx + a;
        int rlx = imax(c->b_lx[idx], c->ilx[i]), rly = imax(c->b_ly[idx], c->ily[i]);
        int rux = imin(c->b_ux[idx], c->iux[i]), ruy = imin(c->b_uy[idx], c->iuy[i]);
        int64_t ov = (rlx >= rux || rly >= ruy) ? 0 : (int64_t)(rux - rlx) * (int64_t)(ruy - rly);
        c->b_nonplace[idx] += (int64_t)((float)ov * c->td);
      }
    }
  }
}

int ro_nesterov_setup(ro_ctx* c, rp_nb_info* info) {
  const rp_params* p = &c->p;
  int ng_i = c->n_place;
  /* fillers, initFillerGCells src/nesterovBase.cpp:1066-1155 */
  int* dxs = NEW(int, ng_i);
  int* dys = NEW(int, ng_i);
  for (int k = 0; k < ng_i; k++) {
    int i = c->place[k];
    dxs[k] = c->iux[i] - c->ilx[i];
    dys[k] = c->iuy[i] - c->ily[i];
  }
  qsort(dxs, (size_t)ng_i, sizeof(int), cmp_int);
  qsort(dys, (size_t)ng_i, sizeof(int), cmp_int);
  int minIdx = (int)((double)ng_i * 0.05);
  int maxIdx = (int)((double)ng_i * 0.95);
  int64_t dxSum = 0, dySum = 0;
  for (int k = minIdx; k < maxIdx; k++) {
    dxSum += dxs[k];
    dySum += dys[k];
  }
  free(dxs);
  free(dys);
  int avgDx = (maxIdx - minIdx) > 0 ? (int)(dxSum / (maxIdx - minIdx)) : 0;
  int avgDy = (maxIdx - minIdx) > 0 ? (int)(dySum / (maxIdx - minIdx)) : 0;
  int64_t coreArea = (int64_t)(c->core_ux - c->core_lx) * (int64_t)(c->core_uy - c->core_ly);
  int64_t whiteSpace = coreArea - c->nonplace_area;
  int64_t movable = (int64_t)((float)whiteSpace * p->target_density);
  int64_t totalFiller = movable - c->std_area - (int64_t)((float)c->macro_area * p->target_density);
  if (totalFiller < 0) return 1;
  int fillerCnt = (avgDx * avgDy) > 0 ? (int)(totalFiller / (int64_t)(avgDx * avgDy)) : 0;
  c->filler_dx = avgDx;
  c->filler_dy = avgDy;

  c->n_gi = ng_i;
  c->n_gf = fillerCnt;
  c->n_g = ng_i + fillerCnt;
  int ng = c->n_g;
  c->g_inst = NEW(int, ng);
  c->g_macro = NEW(uint8_t, ng);
  c->g_lx = NEW(int, ng); c->g_ly = NEW(int, ng); c->g_ux = NEW(int, ng); c->g_uy = NEW(int, ng);
  c->g_dlx = NEW(int, ng); c->g_dly = NEW(int, ng); c->g_dux = NEW(int, ng); c->g_duy = NEW(int, ng);
  c->g_scale = NEW(float, ng);
  for (int k = 0; k < ng_i; k++) {
    int i = c->place[k];
    c->g_inst[k] = i;
    c->g_lx[k] = c->g_dlx[k] = c->ilx[i];
    c->g_ly[k] = c->g_dly[k] = c->ily[i];
    c->g_ux[k] = c->g_dux[k] = c->iux[i];
    c->g_uy[k] = c->g_duy[k] = c->iuy[i];
    c->g_macro[k] = (c->iuy[i] - c->ily[i]) > c->site_h * 6; /* :924 */
  }
  mt_state mt;
  mt_seed(&mt, 0);
  int cdx = c->core_ux - c->core_lx, cdy = c->core_uy - c->core_ly;
  for (int k = 0; k < fillerCnt; k++) {
    uint64_t rx = mt_next(&mt);
    uint64_t ry = mt_next(&mt);
    int cx = (int)(rx % (uint64_t)(int64_t)cdx + (uint64_t)(int64_t)c->core_lx);
    int cy = (int)(ry % (uint64_t)(int64_t)cdy + (uint64_t)(int64_t)c->core_ly);
    int g = ng_i + k;
    c->g_inst[g] = -1;
    c->g_lx[g] = c->g_dlx[g] = cx - avgDx / 2; /* GCell(cx,cy,dx,dy) :58-65 */
    c->g_ly[g] = c->g_dly[g] = cy - avgDy / 2;
    c->g_ux[g] = c->g_dux[g] = cx + avgDx / 2;
    c->g_uy[g] = c->g_duy[g] = cy + avgDy / 2;
  }
  /* GPins / GNets, src/nesterovBase.cpp:937-1001 */
  int np = c->n_pin;
  c->gp_cell = NEW(int, np);
  c->gp_offx = NEW(int, np);
  c->gp_offy = NEW(int, np);
  c->gp_cx = NEW(int, np);
  c->gp_cy = NEW(int, np);
  c->gp_exp = NEW(float, 4 * np);
  c->gp_has = NEW(uint8_t, np);
  for (int q = 0; q < np; q++) {
    int i = c->pin_inst[q];
    c->gp_cell[q] = (i >= 0) ? c->inst2place[i] : -1;
    c->gp_offx[q] = c->pin_offx[q];
    c->gp_offy[q] = c->pin_offy[q];
    c->gp_cx[q] = c->pin_cx[q];
    c->gp_cy[q] = c->pin_cy[q];
  }
  c->g_pstart = NEW(int, ng + 1);
  int tot = 0;
  for (int g = 0; g < ng; g++) {
    c->g_pstart[g] = tot;
    if (g < ng_i) {
      int i = c->g_inst[g];
      tot += c->ipin_start[i + 1] - c->ipin_start[i];
    }
  }
  c->g_pstart[ng] = tot;
  c->g_pins = NEW(int, tot);
  for (int g = 0; g < ng_i; g++) {
    int i = c->g_inst[g];
    memcpy(c->g_pins + c->g_pstart[g], c->ipins + c->ipin_start[i],
           sizeof(int) * (size_t)(c->ipin_start[i + 1] - c->ipin_start[i]));
  }
  c->gn_box = NEW(int, 4 * c->n_net);
  c->gn_sum = NEW(float, 8 * c->n_net);

  /* BinGrid::initBins, src/nesterovBase.cpp:631-717 */
  c->td = p->target_density;
  int64_t totalBinArea = (int64_t)cdx * (int64_t)cdy;
  int64_t avgArea = ng_i > 0 ? c->place_area / (int64_t)ng_i : 0;
  int64_t idealBinArea = (int64_t)roundf((float)avgArea / c->td);
  int idealBinCnt = idealBinArea > 0 ? (int)(totalBinArea / idealBinArea) : 0;
  int found = 2;
  for (found = 2; found <= 1024; found *= 2) {
    if (found * found <= idealBinCnt && 4 * found * found > idealBinCnt) break;
  }
  c->nbx = p->bin_grid_cnt_x ? p->bin_grid_cnt_x : found;
  c->nby = p->bin_grid_cnt_y ? p->bin_grid_cnt_y : found;
  c->bsx = (int)ceilf((float)cdx / (float)c->nbx);
  c->bsy = (int)ceilf((float)cdy / (float)c->nby);
  int nb = c->nbx * c->nby;
  c->b_lx = NEW(int, nb); c->b_ly = NEW(int, nb); c->b_ux = NEW(int, nb); c->b_uy = NEW(int, nb);
  {
    int x = c->core_lx, y = c->core_ly, ix = 0, iy = 0;
    for (int b = 0; b < nb; b++) {
      int sx = (x + c->bsx > c->core_ux) ? c->core_ux - x : c->bsx;
      int sy = (y + c->bsy > c->core_uy) ? c->core_uy - y : c->bsy;
      c->b_lx[b] = x;
      c->b_ly[b] = y;
      c->b_ux[b] = x + sx;
      c->b_uy[b] = y + sy;
      x += c->bsx;
      ix++;
      if (x >= c->core_ux) {
        y += c->bsy;
        x = c->core_lx;
        iy++;
        ix = 0;
      }
    }
  }
  c->b_nonplace = NEW(int64_t, nb);
  c->b_inst = NEW(int64_t, nb);
  c->b_filler = NEW(int64_t, nb);
  c->b_density = NEW(float, nb);
  c->b_phi = NEW(float, nb);
  c->b_ex = NEW(float, nb);
  c->b_ey = NEW(float, nb);
  bins_nonplace(c);

  /* density size and scale, src/nesterovBase.cpp:1033-1061 */
  for (int g = 0; g < ng; g++) {
    int gdx = c->g_ux[g] - c->g_lx[g], gdy = c->g_uy[g] - c->g_ly[g];
    float scaleX, scaleY, dsx, dsy;
    if ((long double)gdx < RO_SQRT2 * (long double)c->bsx) {
      scaleX = (float)gdx / (float)(RO_SQRT2 * (long double)c->bsx);
      dsx = (float)(RO_SQRT2 * (long double)(float)c->bsx);
    } else {
      scaleX = 1.0f;
      dsx = (float)gdx;
    }
    if ((long double)gdy < RO_SQRT2 * (long double)c->bsy) {
      scaleY = (float)gdy / (float)(RO_SQRT2 * (long double)c->bsy);
      dsy = (float)(RO_SQRT2 * (long double)(float)c->bsy);
    } else {
      scaleY = 1.0f;
      dsy = (float)gdy;
    }
    gcell_set_density_size(c, g, (int)dsx, (int)dsy);
    c->g_scale[g] = scaleX * scaleY;
  }
  c->nb_ready = 1;
  if (info) {
    memset(info, 0, sizeof(*info));
    info->num_gcells = ng;
    info->num_fillers = fillerCnt;
    info->num_place_insts = ng_i;
    info->num_gnets = c->n_net;
    info->num_gpins = c->n_pin;
    info->bin_cnt_x = c->nbx;
    info->bin_cnt_y = c->nby;
    info->bin_size_x = c->bsx;
    info->bin_size_y = c->bsy;
    info->core_lx = c->core_lx;
    info->core_ly = c->core_ly;
    info->core_ux = c->core_ux;
    info->core_uy = c->core_uy;
    info->filler_dx = avgDx;
    info->filler_dy = avgDy;
    info->std_insts_area = c->std_area;
    info->macro_insts_area = c->macro_area;
    info->place_insts_area = c->place_area;
    info->nonplace_insts_area = c->nonplace_area;
    info->target_density = c->td;
    info->core_area = coreArea;
    info->white_space_area = whiteSpace;
    info->movable_area = movable;
    info->total_filler_area = totalFiller;
    info->average_place_inst_area = avgArea;
    info->ideal_bin_area = idealBinArea;
    info->ideal_bin_count = idealBinCnt;
    info->total_bin_area = totalBinArea;
  }
  /* updateDensityCoordiLayoutInside for every gcell, NesterovPlace::init
   * src/nesterovPlace.cpp:86-93 -> src/nesterovBase.cpp:1245-1268 */
  for (int g = 0; g < ng; g++) {
    float tlx = (float)c->g_dlx[g], tly = (float)c->g_dly[g];
    int ddx = c->g_dux[g] - c->g_dlx[g], ddy = c->g_duy[g] - c->g_dly[g];
    if (tlx < (float)c->core_lx) tlx = (float)c->core_lx;
    if (tly < (float)c->core_ly) tly = (float)c->core_ly;
    if (tlx + (float)ddx > (float)c->core_ux) tlx = (float)(c->core_ux - ddx);
    if (tly + (float)ddy > (float)c->core_uy) tly = (float)(c->core_uy - ddy);
    /* setDensityLocation(int, int), src/nesterovBase.cpp:151-162 */
    int nlx = (int)tlx, nly = (int)tly;
    c->g_dux[g] = nlx + ddx;
    c->g_duy[g] = nly + ddy;
    c->g_dlx[g] = nlx;
    c->g_dly[g] = nly;
    gcell_update_pins(c, g);
  }
  return 0;
}

int ro_get_gcells(ro_ctx* c, int32_t* dhx, int32_t* dhy, float* scale, int32_t* dx, int32_t* dy,
                  float* icx, float* icy, uint8_t* kind) {
  if (!c->nb_ready) return RP_ERR_STATE;
  for (int g = 0; g < c->n_g; g++) {
    if (dhx) dhx[g] = (c->g_dux[g] - c->g_dlx[g]) / 2;
    if (dhy) dhy[g] = (c->g_duy[g] - c->g_dly[g]) / 2;
    if (scale) scale[g] = c->g_scale[g];
    if (dx) dx[g] = c->g_ux[g] - c->g_lx[g];
    if (dy) dy[g] = c->g_uy[g] - c->g_ly[g];
    if (icx) icx[g] = (float)((c->g_dux[g] + c->g_dlx[g]) / 2);
    if (icy) icy[g] = (float)((c->g_duy[g] + c->g_dly[g]) / 2);
    if (kind) kind[g] = c->g_inst[g] < 0 ? 2 : (c->g_macro[g] ? 1 : 0);
  }
  return 0;
}

int ro_get_nonplace(ro_ctx* c, int64_t* out) {
  if (!c->nb_ready) return RP_ERR_STATE;
  memcpy(out, c->b_nonplace, sizeof(int64_t) * (size_t)c->nbx * c->nby);
  return 0;
}

static void density_minmax(ro_ctx* c, int g, int* x0, int* x1, int* y0, int* y1) {
  /* getDensityMinMaxIdxX/Y, src/nesterovBase.cpp:822-841 (no clamp) */
  int r;
  *x0 = (c->g_dlx[g] - c->core_lx) / c->bsx;
  r = c->g_dux[g] - c->core_lx;
  *x1 = ((r >= c->bsx ? r % c->bsx : r) == 0) ? r / c->bsx : r / c->bsx + 1;
  *y0 = (c->g_dly[g] - c->core_ly) / c->bsy;
  r = c->g_duy[g] - c->core_ly;
  *y1 = ((r >= c->bsy ? r % c->bsy : r) == 0) ? r / c->bsy : r / c->bsy + 1;
}

static float overlap_density_area(ro_ctx* c, int b, int g) {
  /* getOverlapDensityArea, src/nesterovBase.cpp:1533-1547 */
  int rlx = imax(c->b_lx[b], c->g_dlx[g]), rly = imax(c->b_ly[b], c->g_dly[g]);
  int rux = imin(c->b_ux[b], c->g_dux[g]), ruy = imin(c->b_uy[b], c->g_duy[g]);
  if (rlx >= rux || rly >= ruy) return 0;
  return (float)(rux - rlx) * (float)(ruy - rly);
}

static void bins_scatter(ro_ctx* c) {
  /* BinGrid::updateBinsGCellDensityArea, src/nesterovBase.cpp:745-819 */
  int nb = c->nbx * c->nby;
  for (int b = 0; b < nb; b++) c->b_inst[b] = c->b_filler[b] = 0;
  for (int g = 0; g < c->n_g; g++) {
    int x0, x1, y0, y1;
    density_minmax(c, g, &x0, &x1, &y0, &y1);
    for (int i = x0; i < x1; i++) {
      for (int j = y0; j < y1; j++) {
        int b = j * c->nbx + i;
        if (c->g_inst[g] >= 0) {
          if (c->g_macro[g])
            c->b_inst[b] += (int64_t)(overlap_density_area(c, b, g) * c->g_scale[g] * c->td);
          else
            c->b_inst[b] += (int64_t)(overlap_density_area(c, b, g) * c->g_scale[g]);
        } else {
          c->b_filler[b] += (int64_t)(overlap_density_area(c, b, g) * c->g_scale[g]);
        }
      }
    }
  }
  c->overflow_area = 0;
  for (int b = 0; b < nb; b++) {
    int64_t binArea = (int64_t)(c->b_ux[b] - c->b_lx[b]) * (int64_t)(c->b_uy[b] - c->b_ly[b]);
    float at = (float)binArea * c->td;
    c->b_density[b] = ((float)c->b_inst[b] + (float)c->b_filler[b] + (float)c->b_nonplace[b]) / at;
    float ov = (float)c->b_inst[b] + (float)c->b_nonplace[b] - at;
    ov = (0.0f < ov) ? ov : 0.0f; /* std::max(0.0f, ov) */
    c->overflow_area = (int64_t)((float)c->overflow_area + ov);
  }
}

static void update_density_center(ro_ctx* c, const float* xy) {
  /* updateGCellDensityCenterLocation, src/nesterovBase.cpp:1198-1207 */
  for (int g = 0; g < c->n_g; g++) gcell_set_density_center(c, g, (int)xy[2 * g], (int)xy[2 * g + 1]);
  bins_scatter(c);
}

static void density_force_bin(ro_ctx* c) {
  /* updateDensityForceBin, src/nesterovBase.cpp:1481-1506 */
  c->pfn(c->puser, c->nbx, c->nby, c->bsx, c->bsy, c->b_density, c->b_phi, c->b_ex, c->b_ey);
  c->sum_phi = 0;
  for (int b = 0; b < c->nbx * c->nby; b++)
    c->sum_phi += c->b_phi[b] * (float)(c->b_nonplace[b] + c->b_inst[b] + c->b_filler[b]);
}

static void net_update_box(ro_ctx* c, int e) {
  /* GNet::updateBox, src/nesterovBase.cpp:273-284 */
  int lx = INT_MAX, ly = INT_MAX, ux = INT_MIN, uy = INT_MIN;
  for (int q = c->net_start[e]; q < c->net_start[e + 1]; q++) {
    lx = imin(c->gp_cx[q], lx);
    ly = imin(c->gp_cy[q], ly);
    ux = imax(c->gp_cx[q], ux);
    uy = imax(c->gp_cy[q], uy);
  }
  c->gn_box[4 * e + 0] = lx;
  c->gn_box[4 * e + 1] = ly;
  c->gn_box[4 * e + 2] = ux;
  c->gn_box[4 * e + 3] = uy;
}

static void wa_force(ro_ctx* c, float cx_, float cy_) {
  /* updateWireLengthForceWA, src/nesterovBase.cpp:1304-1360 */
  float bar = c->p.min_wire_length_force_bar;
  memset(c->gn_sum, 0, sizeof(float) * 8 * (size_t)c->n_net);
  memset(c->gp_exp, 0, sizeof(float) * 4 * (size_t)c->n_pin);
  memset(c->gp_has, 0, (size_t)c->n_pin);
  for (int e = 0; e < c->n_net; e++) {
    net_update_box(c, e);
    int* bx = c->gn_box + 4 * e;
    float* s = c->gn_sum + 8 * e;
    for (int q = c->net_start[e]; q < c->net_start[e + 1]; q++) {
      int px = c->gp_cx[q], py = c->gp_cy[q];
      float eMinX = (float)(bx[0] - px) * cx_;
      float eMaxX = (float)(px - bx[2]) * cx_;
      float eMinY = (float)(bx[1] - py) * cy_;
      float eMaxY = (float)(py - bx[3]) * cy_;
      float* pe = c->gp_exp + 4 * q;
      if (eMinX > bar) {
        pe[0] = ro_fast_exp(eMinX);
        c->gp_has[q] |= 1;
        s[0] += pe[0];
        s[1] += (float)px * pe[0];
      }
      if (eMaxX > bar) {
        pe[1] = ro_fast_exp(eMaxX);
        c->gp_has[q] |= 2;
        s[2] += pe[1];
        s[3] += (float)px * pe[1];
      }
      if (eMinY > bar) {
        pe[2] = ro_fast_exp(eMinY);
        c->gp_has[q] |= 4;
        s[4] += pe[2];
        s[5] += (float)py * pe[2];
      }
      if (eMaxY > bar) {
        pe[3] = ro_fast_exp(eMaxY);
        c->gp_has[q] |= 8;
        s[6] += pe[3];
        s[7] += (float)py * pe[3];
      }
    }
  }
}

static void wl_grad_pin(ro_ctx* c, int q, float cx_, float cy_, float* gx, float* gy) {
  /* getWireLengthGradientPinWA, src/nesterovBase.cpp:1383-1440 */
  float gMinX = 0, gMinY = 0, gMaxX = 0, gMaxY = 0;
  int e = -1;
  /* pin q belongs to the net whose range contains it (pins are net-major) */
  {
    int lo = 0, hi = c->n_net - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) / 2;
      if (c->net_start[mid] <= q)
        lo = mid;
      else
        hi = mid - 1;
    }
    e = lo;
  }
  const float* s = c->gn_sum + 8 * e;
  const float* pe = c->gp_exp + 4 * q;
  int px = c->gp_cx[q], py = c->gp_cy[q];
  if (c->gp_has[q] & 1) {
    float S = s[0], XS = s[1];
    gMinX = (float)(((double)S * ((double)pe[0] * (1.0 - (double)(cx_ * (float)px))) +
                     (double)(cx_ * pe[0] * XS)) /
                    (double)(S * S));
  }
  if (c->gp_has[q] & 2) {
    float S = s[2], XS = s[3];
    gMaxX = (float)(((double)S * ((double)pe[1] * (1.0 + (double)(cx_ * (float)px))) -
                     (double)(cx_ * pe[1] * XS)) /
                    (double)(S * S));
  }
  if (c->gp_has[q] & 4) {
    float S = s[4], XS = s[5];
    gMinY = (float)(((double)S * ((double)pe[2] * (1.0 - (double)(cy_ * (float)py))) +
                     (double)(cy_ * pe[2] * XS)) /
                    (double)(S * S));
  }
  if (c->gp_has[q] & 8) {
    float S = s[6], XS = s[7];
    gMaxY = (float)(((double)S * ((double)pe[3] * (1.0 + (double)(cy_ * (float)py))) -
                     (double)(cy_ * pe[3] * XS)) /
                    (double)(S * S));
  }
  *gx = gMinX - gMaxX;
  *gy = gMinY - gMaxY;
}

static void density_grad(ro_ctx* c, int g, float* fx, float* fy) {
  /* getDensityGradient, src/nesterovBase.cpp:1458-1478 */
  int x0, x1, y0, y1;
  density_minmax(c, g, &x0, &x1, &y0, &y1);
  float ex = 0, ey = 0;
  for (int i = x0; i < x1; i++) {
    for (int j = y0; j < y1; j++) {
      int b = j * c->nbx + i;
      float ov = overlap_density_area(c, b, g) * c->g_scale[g];
      ex += ov * c->b_ex[b];
      ey += ov * c->b_ey[b];
    }
  }
  *fx = ex;
  *fy = ey;
}

static void update_gradients(ro_ctx* c, float* sumG, float* wlG, float* dG, double* sums_out) {
  /* NesterovPlace::updateGradients, src/nesterovPlace.cpp:206-273 */
  c->wlGradSum = 0;
  c->densGradSum = 0;
  float gradSum = 0;
  for (int g = 0; g < c->n_g; g++) {
    float wx = 0, wy = 0;
    for (int k = c->g_pstart[g]; k < c->g_pstart[g + 1]; k++) {
      float tx, ty;
      wl_grad_pin(c, c->g_pins[k], c->wlCoefX, c->wlCoefY, &tx, &ty);
      wx += tx;
      wy += ty;
    }
    float dx, dy;
    density_grad(c, g, &dx, &dy);
    wlG[2 * g] = wx;
    wlG[2 * g + 1] = wy;
    dG[2 * g] = dx;
    dG[2 * g + 1] = dy;
    c->wlGradSum += fabsf(wx);
    c->wlGradSum += fabsf(wy);
    c->densGradSum += fabsf(dx);
    c->densGradSum += fabsf(dy);
    float sx = wx + c->densityPenalty * dx;
    float sy = wy + c->densityPenalty * dy;
    float npin = (float)(c->g_pstart[g + 1] - c->g_pstart[g]);
    float area = (float)(c->g_ux[g] - c->g_lx[g]) * (float)(c->g_uy[g] - c->g_ly[g]);
    float px = npin + c->densityPenalty * area;
    float py = npin + c->densityPenalty * area;
    if (px <= c->p.min_preconditioner) px = c->p.min_preconditioner;
    if (py <= c->p.min_preconditioner) py = c->p.min_preconditioner;
    sx /= px;
    sy /= py;
    sumG[2 * g] = sx;
    sumG[2 * g + 1] = sy;
    gradSum += fabsf(sx) + fabsf(sy);
  }
  if (sums_out) {
    sums_out[0] = c->wlGradSum;
    sums_out[1] = c->densGradSum;
    sums_out[2] = gradSum;
  }
  if (isnan(c->wlGradSum) || isinf(c->wlGradSum) || isnan(c->densGradSum) || isinf(c->densGradSum))
    c->isDiverged = 1;
}

static int64_t nb_hpwl(ro_ctx* c) {
  /* NesterovBase::getHpwl, src/nesterovBase.cpp:1508-1516 */
  int64_t h = 0;
  for (int e = 0; e < c->n_net; e++) {
    net_update_box(c, e);
    h += (int64_t)((c->gn_box[4 * e + 2] - c->gn_box[4 * e]) + (c->gn_box[4 * e + 3] - c->gn_box[4 * e + 1]));
  }
  return h;
}

int ro_eval_gradient(ro_ctx* c, const float* cx, const float* cy, float lam, float wlx, float wly,
                     rp_eval_out* o) {
  if (!c->nb_ready) return RP_ERR_STATE;
  int ng = c->n_g, nb = c->nbx * c->nby;
  float* xy = NEW(float, 2 * ng);
  for (int g = 0; g < ng; g++) {
    xy[2 * g] = cx[g];
    xy[2 * g + 1] = cy[g];
  }
  update_density_center(c, xy);
  density_force_bin(c);
  float savedL = c->densityPenalty, savedX = c->wlCoefX, savedY = c->wlCoefY;
  int savedDiv = c->isDiverged;
  c->densityPenalty = lam;
  c->wlCoefX = wlx;
  c->wlCoefY = wly;
  wa_force(c, wlx, wly);
  float* sg = NEW(float, 2 * ng);
  float* wg = NEW(float, 2 * ng);
  float* dg = NEW(float, 2 * ng);
  double sums[3];
  update_gradients(c, sg, wg, dg, sums);
  if (o->bin_inst) memcpy(o->bin_inst, c->b_inst, sizeof(int64_t) * nb);
  if (o->bin_filler) memcpy(o->bin_filler, c->b_filler, sizeof(int64_t) * nb);
  if (o->density) memcpy(o->density, c->b_density, sizeof(float) * nb);
  if (o->phi) memcpy(o->phi, c->b_phi, sizeof(float) * nb);
  if (o->ex) memcpy(o->ex, c->b_ex, sizeof(float) * nb);
  if (o->ey) memcpy(o->ey, c->b_ey, sizeof(float) * nb);
  if (o->overflow_area) *o->overflow_area = c->overflow_area;
  if (o->net_sums) memcpy(o->net_sums, c->gn_sum, sizeof(float) * 8 * c->n_net);
  if (o->net_box) memcpy(o->net_box, c->gn_box, sizeof(int) * 4 * c->n_net);
  if (o->wl_grad) memcpy(o->wl_grad, wg, sizeof(float) * 2 * ng);
  if (o->dens_grad) memcpy(o->dens_grad, dg, sizeof(float) * 2 * ng);
  if (o->sum_grad) memcpy(o->sum_grad, sg, sizeof(float) * 2 * ng);
  if (o->sums) memcpy(o->sums, sums, sizeof(double) * 3);
  if (o->hpwl) *o->hpwl = nb_hpwl(c);
  c->densityPenalty = savedL;
  c->wlCoefX = savedX;
  c->wlCoefY = savedY;
  c->isDiverged = savedDiv;
  free(xy);
  free(sg);
  free(wg);
  free(dg);
  return 0;
}

int ro_eval_poisson(ro_ctx* c, const float* density, float* phi, float* ex, float* ey) {
  if (!c->nb_ready) return RP_ERR_STATE;
  c->pfn(c->puser, c->nbx, c->nby, c->bsx, c->bsy, density, phi, ex, ey);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Poisson: own FP64 FFT-based transforms (Makhoul 1980)              */
/* ------------------------------------------------------------------ */
static void fft_c(double* re, double* im, int n, int inverse) {
  /* iterative radix-2 decimation in time, unnormalised; sign -1 forward */
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    double ang = 2.0 * M_PI / (double)len * (inverse ? 1.0 : -1.0);
    for (int i = 0; i < n; i += len) {
      for (int k = 0; k < len / 2; k++) {
        double wr = cos(ang * k), wi = sin(ang * k);
        int a = i + k, b = i + k + len / 2;
        double xr = re[b] * wr - im[b] * wi;
        double xi = re[b] * wi + im[b] * wr;
        re[b] = re[a] - xr;
        im[b] = im[a] - xi;
        re[a] += xr;
        im[a] += xi;
      }
    }
  }
}

/* X[k] = sum_n x[n] cos(pi (n+1/2) k / N)  (Ooura ddct isgn=-1) */
static void dct2_64(double* x, int n, double* re, double* im) {
  for (int k = 0; k < n / 2; k++) {
    re[k] = x[2 * k];
    re[n - 1 - k] = x[2 * k + 1];
  }
  for (int k = 0; k < n; k++) im[k] = 0.0;
  fft_c(re, im, n, 0);
  for (int k = 0; k < n; k++) {
    double a = -M_PI * k / (2.0 * n);
    x[k] = re[k] * cos(a) - im[k] * sin(a);
  }
}

/* y[k] = sum_j a[j] cos(pi j (k+1/2) / N)  (Ooura ddct isgn=+1) */
static void dct3_64(double* a, int n, double* re, double* im) {
  re[0] = 2.0 * a[0];
  im[0] = 0.0;
  for (int k = 1; k < n; k++) {
    double ang = M_PI * k / (2.0 * n);
    double ar = a[k], ai = -a[n - k];
    re[k] = ar * cos(ang) - ai * sin(ang);
    im[k] = ar * sin(ang) + ai * cos(ang);
  }
  fft_c(re, im, n, 1);
  for (int k = 0; k < n / 2; k++) {
    a[2 * k] = 0.5 * re[k];
    a[2 * k + 1] = 0.5 * re[n - 1 - k];
  }
}

/* S[k] = sum_{j=1}^{N} A[j] sin(pi j (k+1/2)/N), A[N] = a[0]  (ddst isgn=+1) */
static void dst3_64(double* a, int n, double* re, double* im, double* tmp) {
  tmp[0] = a[0];
  for (int j = 1; j < n; j++) tmp[j] = a[n - j];
  dct3_64(tmp, n, re, im);
  for (int k = 0; k < n; k++) a[k] = (k & 1) ? -tmp[k] : tmp[k];
}

/* FFT::doFFT (src/fft.cpp:112-177) with FP64 transforms; the float
 * rounding between stages follows the reference (binDensity_ etc. are
 * float arrays). */
void ro_poisson_fft64(void* user, int nx, int ny, int bsx, int bsy, const float* rho, float* phi,
                      float* ex, float* ey) {
  (void)user;
  int nm = nx > ny ? nx : ny;
  double* re = (double*)malloc(sizeof(double) * nm);
  double* im = (double*)malloc(sizeof(double) * nm);
  double* tmp = (double*)malloc(sizeof(double) * nm);
  double* line = (double*)malloc(sizeof(double) * nm);
  double* A = (double*)malloc(sizeof(double) * (size_t)nx * ny); /* [i][j] x-major */
  float* F = (float*)malloc(sizeof(float) * (size_t)nx * ny);
  float* P = (float*)malloc(sizeof(float) * (size_t)nx * ny);
  float* EX = (float*)malloc(sizeof(float) * (size_t)nx * ny);
  float* EY = (float*)malloc(sizeof(float) * (size_t)nx * ny);
  float* wx = (float*)malloc(sizeof(float) * nx);
  float* wy = (float*)malloc(sizeof(float) * ny);
  /* FFT::init twiddles, src/fft.cpp:78-90 */
  for (int i = 0; i < nx; i++) wx[i] = (float)(RO_PI * (long double)(float)i / (long double)(float)nx);
  for (int j = 0; j < ny; j++)
    wy[j] = (float)(RO_PI * (long double)(float)j / (long double)(float)ny * (long double)(float)bsy /
                    (long double)(float)bsx);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) A[(size_t)i * ny + j] = rho[(size_t)j * nx + i];
  /* ddct2d(-1): along j for every i, then along i */
  for (int i = 0; i < nx; i++) dct2_64(A + (size_t)i * ny, ny, re, im);
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) line[i] = A[(size_t)i * ny + j];
    dct2_64(line, nx, re, im);
    for (int i = 0; i < nx; i++) A[(size_t)i * ny + j] = line[i];
  }
  for (size_t k = 0; k < (size_t)nx * ny; k++) F[k] = (float)A[k];
  for (int i = 0; i < nx; i++) F[(size_t)i * ny] *= 0.5f;
  for (int j = 0; j < ny; j++) F[j] *= 0.5f;
  for (size_t k = 0; k < (size_t)nx * ny; k++) F[k] = (float)((double)F[k] * (4.0 / nx / ny));
  for (int i = 0; i < nx; i++) {
    float wxi = wx[i], wx2 = wx[i] * wx[i];
    for (int j = 0; j < ny; j++) {
      float wyj = wy[j], wy2 = wy[j] * wy[j];
      size_t k = (size_t)i * ny + j;
      float d = F[k], ph = 0, exv = 0, eyv = 0;
      if (!(i == 0 && j == 0)) {
        ph = d / (wx2 + wy2);
        exv = ph * wxi;
        eyv = ph * wyj;
      }
      P[k] = ph;
      EX[k] = exv;
      EY[k] = eyv;
    }
  }
  /* inverse: phi = ddct2d(+1); ex = ddsct2d(+1); ey = ddcst2d(+1) */
  for (int pass = 0; pass < 3; pass++) {
    float* src = pass == 0 ? P : (pass == 1 ? EX : EY);
    for (size_t k = 0; k < (size_t)nx * ny; k++) A[k] = src[k];
    for (int i = 0; i < nx; i++) {
      if (pass == 2)
        dst3_64(A + (size_t)i * ny, ny, re, im, tmp);
      else
        dct3_64(A + (size_t)i * ny, ny, re, im);
    }
    for (int j = 0; j < ny; j++) {
      for (int i = 0; i < nx; i++) line[i] = A[(size_t)i * ny + j];
      if (pass == 1)
        dst3_64(line, nx, re, im, tmp);
      else
        dct3_64(line, nx, re, im);
      for (int i = 0; i < nx; i++) A[(size_t)i * ny + j] = line[i];
    }
    float* dst = pass == 0 ? phi : (pass == 1 ? ex : ey);
    if (dst)
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) dst[(size_t)j * nx + i] = (float)A[(size_t)i * ny + j];
  }
  free(re); free(im); free(tmp); free(line); free(A); free(F); free(P); free(EX); free(EY);
  free(wx); free(wy);
}

/* Closed form of the Poisson solve (FP64 direct separable sums of the
 * reference definitions, src/fftsg2d.cpp:200-333 and src/fft.cpp:117-166),
 * without any intermediate float rounding. */
void ro_poisson_direct64(void* user, int nx, int ny, int bsx, int bsy, const float* rho,
                         double* phi, double* ex, double* ey) {
  (void)user;
  double* T = (double*)calloc((size_t)nx * ny, sizeof(double));
  double* A = (double*)calloc((size_t)nx * ny, sizeof(double));
  /* A[u][v] = sum_i sum_j rho[i][j] cos(pi(i+1/2)u/nx) cos(pi(j+1/2)v/ny) */
  for (int i = 0; i < nx; i++)
    for (int v = 0; v < ny; v++) {
      double s = 0;
      for (int j = 0; j < ny; j++) s += rho[(size_t)j * nx + i] * cos(M_PI * (j + 0.5) * v / ny);
      T[(size_t)i * ny + v] = s;
    }
  for (int u = 0; u < nx; u++)
    for (int v = 0; v < ny; v++) {
      double s = 0;
      for (int i = 0; i < nx; i++) s += T[(size_t)i * ny + v] * cos(M_PI * (i + 0.5) * u / nx);
      double w = 4.0 / nx / ny;
      if (u == 0) w *= 0.5;
      if (v == 0) w *= 0.5;
      A[(size_t)u * ny + v] = s * w;
    }
  double asp = (double)bsy / (double)bsx;
  for (int pass = 0; pass < 3; pass++) {
    double* out = pass == 0 ? phi : (pass == 1 ? ex : ey);
    for (int u = 0; u < nx; u++)
      for (int j = 0; j < ny; j++) {
        double s = 0;
        for (int v = 0; v < ny; v++) {
          if (u == 0 && v == 0) continue;
          double wxu = M_PI * u / nx, wyv = M_PI * v / ny * asp;
          double ph = A[(size_t)u * ny + v] / (wxu * wxu + wyv * wyv);
          double coef = pass == 0 ? ph : (pass == 1 ? ph * wxu : ph * wyv);
          double basis = pass == 2 ? sin(M_PI * v * (j + 0.5) / ny) : cos(M_PI * v * (j + 0.5) / ny);
          s += coef * basis;
        }
        T[(size_t)u * ny + j] = s;
      }
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        double s = 0;
        for (int u = 0; u < nx; u++) {
          double basis = pass == 1 ? sin(M_PI * u * (i + 0.5) / nx) : cos(M_PI * u * (i + 0.5) / nx);
          s += T[(size_t)u * ny + j] * basis;
        }
        out[(size_t)j * nx + i] = s;
      }
  }
  free(T);
  free(A);
}

/* ------------------------------------------------------------------ */
/* NesterovPlace (src/nesterovPlace.cpp)                              */
/* ------------------------------------------------------------------ */
static float get_distance(const float* a, const float* b, int n) {
  /* getDistance, src/nesterovPlace.cpp:598-607 */
  float s = 0.0f;
  for (int i = 0; i < n; i++) {
    s += (a[2 * i] - b[2 * i]) * (a[2 * i] - b[2 * i]);
    s += (a[2 * i + 1] - b[2 * i + 1]) * (a[2 * i + 1] - b[2 * i + 1]);
  }
  return (float)sqrt((double)s / (2.0 * (double)n));
}

static float step_length(ro_ctx* c, const float* pc, const float* pg, const float* cc, const float* cg) {
  /* getStepLength, src/nesterovPlace.cpp:555-571 */
  float cd = get_distance(pc, cc, c->n_g);
  float gd = get_distance(pg, cg, c->n_g);
  return cd / gd;
}

static void update_wl_coef(ro_ctx* c, float overflow) {
  /* updateWireLengthCoef, src/nesterovPlace.cpp:474-490 */
  if (overflow > 1.0)
    c->wlCoefX = c->wlCoefY = 0.1f;
  else if (overflow < 0.1)
    c->wlCoefX = c->wlCoefY = 10.0f;
  else
    c->wlCoefX = c->wlCoefY = (float)(1.0 / pow(10.0, ((double)overflow - 0.1) * 20 / 9.0 - 1.0));
  c->wlCoefX *= c->baseWLCoef;
  c->wlCoefY *= c->baseWLCoef;
}

static float clamp_x(ro_ctx* c, int g, float x) {
  /* getDensityCoordiLayoutInsideX, src/nesterovBase.cpp:1270-1282 */
  int h = (c->g_dux[g] - c->g_dlx[g]) / 2;
  float v = x;
  if (x - (float)h < (float)c->core_lx) v = (float)(c->core_lx + h);
  if (x + (float)h > (float)c->core_ux) v = (float)(c->core_ux - h);
  return v;
}

static float clamp_y(ro_ctx* c, int g, float y) {
  int h = (c->g_duy[g] - c->g_dly[g]) / 2;
  float v = y;
  if (y - (float)h < (float)c->core_ly) v = (float)(c->core_ly + h);
  if (y + (float)h > (float)c->core_uy) v = (float)(c->core_uy - h);
  return v;
}

static float overflow_of(ro_ctx* c) {
  return (float)c->overflow_area / ((float)c->std_area + (float)c->macro_area * c->td);
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void full_eval(ro_ctx* c, const float* xy, float* sumG, float* wlG, float* dG) {
  double t0 = now_s();
  update_density_center(c, xy);
  double t1 = now_s();
  density_force_bin(c);
  double t2 = now_s();
  wa_force(c, c->wlCoefX, c->wlCoefY);
  double t3 = now_s();
  update_gradients(c, sumG, wlG, dG, NULL);
  double t4 = now_s();
  c->t_stage[RO_T_SCATTER] += t1 - t0;
  c->t_stage[RO_T_POISSON] += t2 - t1;
  c->t_stage[RO_T_WA] += t3 - t2;
  c->t_stage[RO_T_GRAD] += t4 - t3;
  c->evals++;
}

int ro_nesterov_init(ro_ctx* c) {
  /* NesterovPlace::init, src/nesterovPlace.cpp:64-175 */
  if (!c->nb_ready) {
    int r = ro_nesterov_setup(c, NULL);
    if (r) return r;
  }
  int ng = c->n_g;
  size_t sz = 2 * (size_t)ng;
#define AL(x) c->x = NEW(float, sz)
  AL(curSLP); AL(curSLPWL); AL(curSLPD); AL(curSLPSum);
  AL(nextSLP); AL(nextSLPWL); AL(nextSLPD); AL(nextSLPSum);
  AL(prevSLP); AL(prevSLPWL); AL(prevSLPD); AL(prevSLPSum);
  AL(curCo); AL(nextCo);
#undef AL
  c->maxPhiCoef = c->p.max_phi_coef;
  c->densityPenalty = 0;
  c->isDiverged = 0;
  c->evals = 0;
  for (int g = 0; g < ng; g++) {
    float x = (float)((c->g_dux[g] + c->g_dlx[g]) / 2);
    float y = (float)((c->g_duy[g] + c->g_dly[g]) / 2);
    c->curSLP[2 * g] = c->prevSLP[2 * g] = c->curCo[2 * g] = x;
    c->curSLP[2 * g + 1] = c->prevSLP[2 * g + 1] = c->curCo[2 * g + 1] = y;
  }
  update_density_center(c, c->curSLP);
  c->prevHpwl = nb_hpwl(c);
  density_force_bin(c);
  c->baseWLCoef = (float)((double)c->p.init_wire_length_coef / ((double)(float)(c->bsx + c->bsy) * 0.5));
  c->sumOverflow = overflow_of(c);
  update_wl_coef(c, c->sumOverflow);
  wa_force(c, c->wlCoefX, c->wlCoefY);
  update_gradients(c, c->curSLPSum, c->curSLPWL, c->curSLPD, NULL);
  c->evals++;
  if (c->isDiverged) return 0;
  /* updateInitialPrevSLPCoordi, src/nesterovPlace.cpp:492-512 */
  for (int g = 0; g < ng; g++) {
    float px = c->curSLP[2 * g] + c->p.initial_prev_coordi_update_coef * c->curSLPSum[2 * g];
    float py = c->curSLP[2 * g + 1] + c->p.initial_prev_coordi_update_coef * c->curSLPSum[2 * g + 1];
    c->prevSLP[2 * g] = clamp_x(c, g, px);
    c->prevSLP[2 * g + 1] = clamp_y(c, g, py);
  }
  full_eval(c, c->prevSLP, c->prevSLPSum, c->prevSLPWL, c->prevSLPD);
  if (c->isDiverged) return 0;
  c->densityPenalty = (c->wlGradSum / c->densGradSum) * c->p.init_density_penalty;
  c->sumOverflow = overflow_of(c);
  c->stepLength = step_length(c, c->prevSLP, c->prevSLPSum, c->curSLP, c->curSLPSum);
  c->np_ready = 1;
  c->curA = 1.0f;
  c->minSumOverflow = 1e30f;
  c->hpwlWithMinSumOverflow = 1e30f;
  c->isMaxPhiCoefChanged = 0;
  c->iter = 0;
  c->finished = 0;
  c->code = 0;
  free(c->rec);
  c->rec = NEW(rp_iter_record, c->p.nesterov_place_max_iter + 1);
  c->n_rec = 0;
  return 0;
}

static void swap_ptr(float** a, float** b) {
  float* t = *a;
  *a = *b;
  *b = t;
}

static void update_next_iter(ro_ctx* c) {
  /* updateNextIter, src/nesterovPlace.cpp:514-553 */
  swap_ptr(&c->prevSLP, &c->curSLP);
  swap_ptr(&c->prevSLPWL, &c->curSLPWL);
  swap_ptr(&c->prevSLPD, &c->curSLPD);
  swap_ptr(&c->prevSLPSum, &c->curSLPSum);
  swap_ptr(&c->curSLP, &c->nextSLP);
  swap_ptr(&c->curSLPWL, &c->nextSLPWL);
  swap_ptr(&c->curSLPD, &c->nextSLPD);
  swap_ptr(&c->curSLPSum, &c->nextSLPSum);
  swap_ptr(&c->curCo, &c->nextCo);
  c->sumOverflow = overflow_of(c);
  update_wl_coef(c, c->sumOverflow);
  int64_t hpwl = nb_hpwl(c);
  float scaled = (float)(hpwl - c->prevHpwl) / c->p.reference_hpwl;
  /* getPhiCoef, src/nesterovPlace.cpp:573-583 */
  float coef = (scaled < 0) ? c->maxPhiCoef
                            : (float)((double)c->maxPhiCoef * pow((double)c->maxPhiCoef, (double)scaled * -1.0));
  if (coef < c->p.min_phi_coef) coef = c->p.min_phi_coef;
  c->prevHpwl = hpwl;
  c->densityPenalty *= coef;
}

int ro_nesterov_iterate(ro_ctx* c, int n_iters, rp_nesterov_stats* st) {
  /* doNesterovPlace loop body, src/nesterovPlace.cpp:315-463 */
  if (!c->np_ready) return RP_ERR_STATE;
  int ng = c->n_g;
  int stop_at = c->iter + n_iters;
  if (stop_at > c->p.nesterov_place_max_iter) stop_at = c->p.nesterov_place_max_iter;
  while (!c->finished && !c->isDiverged && c->iter < stop_at) {
    int i = c->iter;
    float prevA = c->curA;
    c->curA = (float)((1.0 + sqrt(4.0 * (double)prevA * (double)prevA + 1.0)) * 0.5);
    float coeff = (float)(((double)prevA - 1.0) / (double)c->curA);
    int bt;
    for (bt = 0; bt < c->p.max_back_track; bt++) {
      double ts0 = now_s();
      for (int k = 0; k < ng; k++) {
        float nx = c->curSLP[2 * k] + c->stepLength * c->curSLPSum[2 * k];
        float ny = c->curSLP[2 * k + 1] + c->stepLength * c->curSLPSum[2 * k + 1];
        float sx = nx + coeff * (nx - c->curCo[2 * k]);
        float sy = ny + coeff * (ny - c->curCo[2 * k + 1]);
        c->nextCo[2 * k] = clamp_x(c, k, nx);
        c->nextCo[2 * k + 1] = clamp_y(c, k, ny);
        c->nextSLP[2 * k] = clamp_x(c, k, sx);
        c->nextSLP[2 * k + 1] = clamp_y(c, k, sy);
      }
      c->t_stage[RO_T_STEP] += now_s() - ts0;
      full_eval(c, c->nextSLP, c->nextSLPSum, c->nextSLPWL, c->nextSLPD);
      if (c->isDiverged) break;
      double tc0 = now_s();
      float ns = step_length(c, c->curSLP, c->curSLPSum, c->nextSLP, c->nextSLPSum);
      c->t_stage[RO_T_STEPLEN] += now_s() - tc0;
      if ((double)ns > (double)c->stepLength * 0.95) {
        c->stepLength = ns;
        break;
      } else {
        c->stepLength = ns;
      }
    }
    if (!c->isMaxPhiCoefChanged && c->sumOverflow < 0.35f) {
      c->isMaxPhiCoefChanged = 1;
      c->maxPhiCoef = (float)((double)c->maxPhiCoef * 0.99);
    }
    if (c->p.max_back_track == bt) {
      c->code = 3;
      c->isDiverged = 1;
    }
    if (c->isDiverged) break;
    update_next_iter(c);
    rp_iter_record* r = &c->rec[c->n_rec++];
    r->overflow = c->sumOverflow;
    r->hpwl = c->prevHpwl;
    r->step_length = c->stepLength;
    r->density_penalty = c->densityPenalty;
    r->num_back_track = bt + 1;
    c->iter = i + 1;
    if (c->minSumOverflow > c->sumOverflow) {
      c->minSumOverflow = c->sumOverflow;
      c->hpwlWithMinSumOverflow = (float)c->prevHpwl;
    }
    /* float * 1.2f compared with int64 -> both as float */
    if (c->sumOverflow < 0.3f && c->sumOverflow - c->minSumOverflow >= 0.02f &&
        c->hpwlWithMinSumOverflow * 1.2f < (float)c->prevHpwl) {
      c->code = 4;
      c->isDiverged = 1;
      break;
    }
    if (i > 50 && c->sumOverflow <= c->p.target_overflow) {
      c->finished = 1;
      break;
    }
  }
  if (st) {
    st->iterations = c->iter;
    st->evals = c->evals;
    st->finished = c->finished;
    st->code = c->code;
    st->final_overflow = c->sumOverflow;
    st->final_step_length = c->stepLength;
    st->final_density_penalty = c->densityPenalty;
    st->diverged = c->isDiverged;
    st->final_hpwl = c->prevHpwl;
    st->seconds = 0;
  }
  return c->code;
}

static void update_db(ro_ctx* c) {
  /* updateDb, src/nesterovPlace.cpp:585-595 */
  for (int g = 0; g < c->n_gi; g++) {
    int i = c->g_inst[g];
    int dcx = (c->g_dux[g] + c->g_dlx[g]) / 2, dcy = (c->g_duy[g] + c->g_dly[g]) / 2;
    c->dblx[i] = dcx - (c->g_dux[g] - c->g_dlx[g]) / 2;
    c->dbly[i] = dcy - (c->g_duy[g] - c->g_dly[g]) / 2;
  }
}

int ro_nesterov_place(ro_ctx* c, rp_nesterov_stats* st) {
  int r = ro_nesterov_init(c);
  if (r) return r;
  if (c->isDiverged) {
    if (st) memset(st, 0, sizeof(*st));
    if (st) st->code = 2;
    if (st) st->diverged = 1;
    return 2;
  }
  r = ro_nesterov_iterate(c, c->p.nesterov_place_max_iter, st);
  update_db(c);
  /* a NaN divergence breaks the loop with code 0 (src/nesterovPlace.cpp:371-374, 411-413) */
  return c->code;
}

/* The loop state between two outer iterations (members of NesterovPlace,
 * src/nesterovPlace.h:76-127, and the doNesterovPlace locals kept across
 * ro_nesterov_iterate calls), per gcell in the reference's gcell order. */
static void copy_xy(float* dst, const float* src, int ng) {
  if (dst && src) memcpy(dst, src, sizeof(float) * 2 * (size_t)ng);
}

int ro_get_state(ro_ctx* c, rp_nesterov_state* s) {
  if (!c->np_ready) return RP_ERR_STATE;
  int ng = c->n_g;
  copy_xy(s->cur_slp, c->curSLP, ng);
  copy_xy(s->cur_slp_grad, c->curSLPSum, ng);
  copy_xy(s->prev_slp, c->prevSLP, ng);
  copy_xy(s->prev_slp_grad, c->prevSLPSum, ng);
  copy_xy(s->cur_coord, c->curCo, ng);
  s->step_length = c->stepLength;
  s->density_penalty = c->densityPenalty;
  s->wl_coef_x = c->wlCoefX;
  s->wl_coef_y = c->wlCoefY;
  s->base_wl_coef = c->baseWLCoef;
  s->cur_a = c->curA;
  s->sum_overflow = c->sumOverflow;
  s->min_sum_overflow = c->minSumOverflow;
  s->hpwl_with_min_sum_overflow = c->hpwlWithMinSumOverflow;
  s->max_phi_coef = c->maxPhiCoef;
  s->is_max_phi_coef_changed = c->isMaxPhiCoefChanged;
  s->iter = c->iter;
  s->prev_hpwl = c->prevHpwl;
  return 0;
}

int ro_set_state(ro_ctx* c, const rp_nesterov_state* s) {
  if (!c->np_ready) return RP_ERR_STATE;
  int ng = c->n_g;
  copy_xy(c->curSLP, s->cur_slp, ng);
  copy_xy(c->curSLPSum, s->cur_slp_grad, ng);
  copy_xy(c->prevSLP, s->prev_slp, ng);
  copy_xy(c->prevSLPSum, s->prev_slp_grad, ng);
  copy_xy(c->curCo, s->cur_coord, ng);
  c->stepLength = s->step_length;
  c->densityPenalty = s->density_penalty;
  c->wlCoefX = s->wl_coef_x;
  c->wlCoefY = s->wl_coef_y;
  c->baseWLCoef = s->base_wl_coef;
  c->curA = s->cur_a;
  c->sumOverflow = s->sum_overflow;
  c->minSumOverflow = s->min_sum_overflow;
  c->hpwlWithMinSumOverflow = s->hpwl_with_min_sum_overflow;
  c->maxPhiCoef = s->max_phi_coef;
  c->isMaxPhiCoefChanged = s->is_max_phi_coef_changed;
  c->iter = s->iter;
  c->prevHpwl = s->prev_hpwl;
  c->finished = 0;
  c->isDiverged = 0;
  c->code = 0;
  return 0;
}

int ro_get_iter_records(ro_ctx* c, rp_iter_record* out, int32_t cap, int32_t* count) {
  int n = c->n_rec < cap ? c->n_rec : cap;
  if (out && n > 0) memcpy(out, c->rec, sizeof(rp_iter_record) * n);
  *count = c->n_rec;
  return 0;
}

int ro_get_positions(ro_ctx* c, int32_t* lx, int32_t* ly) {
  for (int i = 0; i < c->n_real; i++) {
    lx[i] = c->dblx[i];
    ly[i] = c->dbly[i];
  }
  return 0;
}

int ro_set_positions(ro_ctx* c, const int32_t* lx, const int32_t* ly) {
  for (int i = 0; i < c->n_real; i++) {
    int w = c->iux[i] - c->ilx[i], h = c->iuy[i] - c->ily[i];
    c->ilx[i] = lx[i];
    c->ily[i] = ly[i];
    c->iux[i] = lx[i] + w;
    c->iuy[i] = ly[i] + h;
    c->dblx[i] = lx[i];
    c->dbly[i] = ly[i];
    inst_update_pins(c, i);
  }
  return 0;
}

int ro_get_stage_times(ro_ctx* c, double* out) {
  for (int k = 0; k < RO_STAGES; k++) out[k] = c->t_stage[k];
  return 0;
}

int ro_set_verbose_trace(ro_ctx* c, int on) {
  (void)c;
  (void)on;
  return 0;
}
