/*
 * fcfc_cli.c -- drop-in executables FCFC_2PT_BOX_GPU (periodic box) and
 * FCFC_2PT_GPU (survey geometry) over the C ABI of libfcfc_gpu.
 *
 * Mirrors the command line and configuration keys of FCFC's FCFC_2PT_BOX /
 * FCFC_2PT (src/fcfc/2pt_box/load_conf.c:485-519, src/fcfc/2pt/load_conf.c:
 * 521-558; CLI options override the file, like libcfg's priorities) and the
 * flow of eval_cf (src/fcfc/2pt_box/eval_cf.c:50-430): per PAIR_COUNT read
 * the catalogues, count pairs on the GPU, double auto counts, normalise,
 * write the pair-count file (same binary/ASCII layout), then evaluate the
 * CF_ESTIMATOR expressions, multipoles and projected CFs.
 *
 * Supported subset: ASCII catalogues (CATALOG_TYPE 0) and FITS binary
 * tables (1); POSITION, WEIGHT and SELECTION as the numeric and logical
 * subset of FCFC's expressions (bitwise operators and strings are
 * refused); COORD_CONVERT with the fiducial cosmology or a (z, distance)
 * table (FCFC_2PT).  HDF5 catalogues are refused.  Extra key with
 * no counterpart in the reference: GPU_PRECISION = 0 (FP64 kernel, default)
 * or 1 (FP32 distances with FP64 re-evaluation near bin edges); CLI --gpu-prec.
 * Compiled twice with -DFCFC_CLI_GEOM=0 (box) or 1 (survey).
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/stat.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "fcfc_gpu.h"

#ifndef FCFC_CLI_GEOM
#define FCFC_CLI_GEOM 0
#endif
#if FCFC_CLI_GEOM == 0
#define CODE_NAME "FCFC_2PT_BOX"
#define DEFAULT_CONF "fcfc_2pt_box.conf"
#else
#define CODE_NAME "FCFC_2PT"
#define DEFAULT_CONF "fcfc_2pt.conf"
#endif

#define MAXKEY 64
#ifndef M_PI
#define M_PI 0x1.921fb54442d18p+1
#endif
#define P_ERR(...) fprintf(stderr, "Error: " __VA_ARGS__)

/* ---------------------------------------------------------------------- */
/* configuration store: KEY -> raw value, with a priority (CLI > file)     */
typedef struct { char key[32]; char *val; int prior; } kv_t;
static kv_t kvs[MAXKEY];
static int nkv = 0;

static void kv_set(const char *key, const char *val, int prior) {
  for (int i = 0; i < nkv; i++) {
    if (!strcmp(kvs[i].key, key)) {
      if (prior >= kvs[i].prior) { free(kvs[i].val); kvs[i].val = strdup(val); kvs[i].prior = prior; }
      return;
    }
  }
  if (nkv < MAXKEY) {
    snprintf(kvs[nkv].key, sizeof kvs[nkv].key, "%s", key);
    kvs[nkv].val = strdup(val);
    kvs[nkv].prior = prior;
    nkv++;
  }
}
static const char *kv_get(const char *key) {
  for (int i = 0; i < nkv; i++) if (!strcmp(kvs[i].key, key)) return kvs[i].val;
  return NULL;
}

static char *trim(char *s) {
  while (isspace((unsigned char) *s)) s++;
  char *e = s + strlen(s);
  while (e > s && isspace((unsigned char) e[-1])) *--e = '\0';
  return s;
}

/* options of load_conf.c:485-519 (box) / 2pt/load_conf.c:521-558 */
typedef struct { char s; const char *l; const char *key; } opt_t;
static const opt_t OPTS[] = {
  {'c', "conf", "CONFIG_FILE"}, {'i', "input", "CATALOG"}, {'l', "label", "CATALOG_LABEL"},
  {'T', "type", "CATALOG_TYPE"}, {0, "skip", "ASCII_SKIP"}, {0, "comment", "ASCII_COMMENT"},
  {'f', "formatter", "ASCII_FORMATTER"}, {'x', "position", "POSITION"}, {'w', "weight", "WEIGHT"},
  {'s', "select", "SELECTION"}, {'b', "box", "BOX_SIZE"}, {'S', "data-struct", "DATA_STRUCT"},
  {'B', "bin", "BINNING_SCHEME"}, {'p', "pair", "PAIR_COUNT"}, {'P', "pair-output", "PAIR_COUNT_FILE"},
  {'e', "cf", "CF_ESTIMATOR"}, {'E', "cf-output", "CF_OUTPUT_FILE"}, {'m', "multipole", "MULTIPOLE"},
  {'M', "mp-output", "MULTIPOLE_FILE"}, {'u', "wp", "PROJECTED_CF"}, {'U', "wp-output", "PROJECTED_FILE"},
  {0, "s-file", "SEP_BIN_FILE"}, {0, "s-min", "SEP_BIN_MIN"}, {0, "s-max", "SEP_BIN_MAX"},
  {0, "s-step", "SEP_BIN_SIZE"}, {0, "mu-num", "MU_BIN_NUM"}, {0, "pi-file", "PI_BIN_FILE"},
  {0, "pi-min", "PI_BIN_MIN"}, {0, "pi-max", "PI_BIN_MAX"}, {0, "pi-step", "PI_BIN_SIZE"},
  {'F', "out-format", "OUTPUT_FORMAT"}, {'O', "overwrite", "OVERWRITE"}, {'v', "verbose", "VERBOSE"},
  {0, "convert", "COORD_CONVERT"}, {'d', "omega-m", "OMEGA_M"}, {0, "omega-l", "OMEGA_LAMBDA"},
  {0, "eos-w", "DE_EOS_W"}, {0, "cmvdst-err", "CMVDST_ERR"}, {0, "cmvdst-file", "Z_CMVDST_FILE"},
  {0, "gpu-prec", "GPU_PRECISION"},
  {0, NULL, NULL}
};

static int read_conf_file(const char *fname, int must) {
  FILE *fp = fopen(fname, "r");
  if (!fp) {
    if (must) { P_ERR("cannot open the configuration file `%s'\n", fname); return FCFC_ERR_FILE; }
    return 0;
  }
  char line[8192];
  while (fgets(line, sizeof line, fp)) {
    char *h = strchr(line, '#');
    if (h) *h = '\0';
    char *eq = strchr(line, '=');
    if (!eq) continue;
    *eq = '\0';
    char *k = trim(line), *v = trim(eq + 1);
    if (*k && *v) kv_set(k, v, 1);
  }
  fclose(fp);
  return 0;
}

/* split "[a, b, c]" or "a" into items; quotes removed */
static int split_list(const char *v, char ***out) {
  char *s = strdup(v ? v : "");
  char *p = trim(s);
  size_t n = strlen(p);
  if (n && p[0] == '[' && p[n - 1] == ']') { p[n - 1] = '\0'; p++; }
  char **items = NULL;
  int cnt = 0;
  char *q = p;
  while (1) {
    char *start = q;
    int inq = 0; char qc = 0;
    while (*q && (inq || *q != ',')) {
      if (!inq && (*q == '"' || *q == '\'')) { inq = 1; qc = *q; }
      else if (inq && *q == qc) inq = 0;
      q++;
    }
    char save = *q;
    *q = '\0';
    char *it = trim(start);
    size_t m = strlen(it);
    if (m >= 2 && (it[0] == '"' || it[0] == '\'') && it[m - 1] == it[0]) { it[m - 1] = '\0'; it++; }
    if (*it || save == ',') {
      items = realloc(items, sizeof(char *) * (cnt + 1));
      items[cnt++] = strdup(it);
    }
    if (!save) break;
    q++;
  }
  free(s);
  *out = items;
  return cnt;
}
static void free_list(char **l, int n) { for (int i = 0; i < n; i++) free(l[i]); free(l); }

static int get_bool(const char *key, int def) {
  const char *v = kv_get(key);
  if (!v) return def;
  return (v[0] == 'T' || v[0] == 't' || v[0] == '1' || v[0] == 'Y' || v[0] == 'y');
}
static int get_double(const char *key, double *d) {
  const char *v = kv_get(key);
  if (!v) return 1;
  char *e;
  *d = strtod(v, &e);
  return (e == v);
}
static int get_int(const char *key, int def) {
  const char *v = kv_get(key);
  return v ? atoi(v) : def;
}

/* ---------------------------------------------------------------------- */
/* ASCII catalogues                                                        */
typedef struct { int64_t n; double *x, *y, *z, *w; double sumw; int loaded; } cat_t;

/* ---------------------------------------------------------------------- */
/* Catalogue expressions for POSITION, WEIGHT and SELECTION: the numeric   */
/* and logical subset of libast (src/lib/libast.c, ast_build/ast_eval):   */
/* numbers, columns $n / ${n}, ( ), abs sqrt ln log isfinite, unary - !,   */
/* then binary ** ; * / % ; + - ; < <= > >= ; == != ; && ; || from the     */
/* tightest, all binary operators left-associative and unary ones binding */
/* tighter than any binary one (libast's precedence table).  Evaluated in */
/* double (% is fmod, ** is pow, a comparison gives 1 or 0) for POSITION  */
/* and WEIGHT, whose columns are converted to double as libast converts   */
/* them for a real-typed tree.  SELECTION keeps libast's dynamic types    */
/* (x_eval_sel): integer columns and all-digit literals are long, so / and */
/* % of two of them are C integer operations and ** an integer power;      */
/* mixed long and double operands promote to double; a row is selected   */
/* when the result is true (non-zero).  Bitwise and shift operators and   */
/* string columns are refused.                                             */
enum { X_NUM, X_COL, X_NEG, X_NOT, X_ABS, X_SQRT, X_LN, X_LOG, X_ISFIN, X_POW,
       X_MUL, X_DIV, X_REM, X_ADD, X_SUB, X_LT, X_LE, X_GT, X_GE, X_EQ, X_NE,
       X_AND, X_OR };
typedef struct { int op; int col; double v; long lv; int islong; } xop_t;
typedef struct {
  xop_t *code; int n, cap, depth, maxdepth, maxcol; const char *p, *err;
  char (*names)[72]; int nnames;       /* FITS column names for ${NAME} */
} xprog_t;

static void x_emit(xprog_t *x, int op, int col, double v) {
  if (x->err) return;
  if (x->n == x->cap) {
    x->cap = x->cap ? 2 * x->cap : 16;
    x->code = realloc(x->code, sizeof(xop_t) * x->cap);
  }
  x->code[x->n++] = (xop_t) {op, col, v, 0, 0};
  if (op == X_NUM || op == X_COL) { if (++x->depth > x->maxdepth) x->maxdepth = x->depth; }
  else if (op >= X_POW) x->depth--;
}
static void x_ws(xprog_t *x) { while (isspace((unsigned char) *x->p)) x->p++; }
static void x_expr(xprog_t *x, int minprec);

static void x_unary(xprog_t *x) {
  x_ws(x);
  const char *p = x->p;
  if (*p == '-') { x->p++; x_unary(x); x_emit(x, X_NEG, 0, 0); return; }
  if (*p == '!' && p[1] != '=') { x->p++; x_unary(x); x_emit(x, X_NOT, 0, 0); return; }
  if (*p == '~') { x->err = "bitwise operators are not supported"; return; }
  if (*p == '(') {
    x->p++; x_expr(x, 0); x_ws(x);
    if (*x->p != ')') { if (!x->err) x->err = "missing `)'"; return; }
    x->p++; return;
  }
  if (*p == '$') {
    p++;
    int brace = (*p == '{');
    if (brace) p++;
    char *e;
    long c = strtol(p, &e, 10);
    if (brace && x->names && (e == p || *e != '}')) {
      /* ${NAME}: a FITS column, by name without regard to case
       * (read_fits.c:get_fts_colnum, CASEINSEN) */
      const char *q = strchr(p, '}');
      if (!q || q == p || q - p >= 72) { x->err = "invalid column name"; return; }
      c = 0;
      for (int k = 0; k < x->nnames && !c; k++)
        if ((size_t) (q - p) == strlen(x->names[k]) && !strncasecmp(p, x->names[k], (size_t) (q - p)))
          c = k + 1;
      if (!c) { x->err = "no FITS column of this name"; return; }
      e = (char *) q;
    }
    if (e == p || c < 1 || c > 65535) { x->err = "invalid column reference"; return; }
    if (brace) { if (*e != '}') { x->err = "invalid column reference"; return; } e++; }
    x->p = e;
    if (c > x->maxcol) x->maxcol = (int) c;
    x_emit(x, X_COL, (int) c, 0);
    return;
  }
  if (*p == '"' || *p == '\'') { x->err = "string values are not supported"; return; }
  static const struct { const char *name; int op; } fn[] = {
    {"abs", X_ABS}, {"sqrt", X_SQRT}, {"ln", X_LN}, {"log", X_LOG}, {"isfinite", X_ISFIN}};
  for (size_t k = 0; k < sizeof fn / sizeof fn[0]; k++) {
    size_t l = strlen(fn[k].name);
    if (!strncmp(p, fn[k].name, l)) {
      const char *q = p + l;
      while (isspace((unsigned char) *q)) q++;
      if (*q == '(') {
        x->p = q + 1; x_expr(x, 0); x_ws(x);
        if (*x->p != ')') { if (!x->err) x->err = "missing `)'"; return; }
        x->p++;
        x_emit(x, fn[k].op, 0, 0);
        return;
      }
    }
  }
  char *e;
  double v = strtod(p, &e);
  if (e == p) { x->err = "unexpected token"; return; }
  x->p = e;
  x_emit(x, X_NUM, 0, v);
  /* digits only: a long literal in SELECTION (libast.c:ast_parse_num) */
  int digits = 1;
  for (const char *q = p; q < e; q++) if (!isdigit((unsigned char) *q)) { digits = 0; break; }
  if (digits && !x->err && v < 9.2e18) {
    x->code[x->n - 1].islong = 1;
    x->code[x->n - 1].lv = strtol(p, NULL, 10);
  }
}

/* binary operator at x->p: its op code, precedence and length (0: none) */
static int x_binop(const char *p, int *prec, int *len) {
  static const struct { const char *s; int op, prec; } ops[] = {
    {"**", X_POW, 11}, {"*", X_MUL, 10}, {"/", X_DIV, 10}, {"%", X_REM, 10},
    {"+", X_ADD, 9}, {"-", X_SUB, 9}, {"<=", X_LE, 7}, {">=", X_GE, 7},
    {"<", X_LT, 7}, {">", X_GT, 7}, {"==", X_EQ, 6}, {"!=", X_NE, 6},
    {"&&", X_AND, 2}, {"||", X_OR, 1}};
  if (!strncmp(p, "<<", 2) || !strncmp(p, ">>", 2)) return -1;
  for (size_t k = 0; k < sizeof ops / sizeof ops[0]; k++) {
    size_t l = strlen(ops[k].s);
    if (!strncmp(p, ops[k].s, l)) { *prec = ops[k].prec; *len = (int) l; return ops[k].op; }
  }
  if (*p == '&' || *p == '|' || *p == '^') return -1;
  return 0;
}

static void x_expr(xprog_t *x, int minprec) {
  x_unary(x);
  for (;;) {
    if (x->err) return;
    x_ws(x);
    int prec, len, op = x_binop(x->p, &prec, &len);
    if (op < 0) { x->err = "bitwise and shift operators are not supported"; return; }
    if (op == 0 || prec < minprec) return;
    x->p += len;
    x_expr(x, prec + 1);                 /* left-associative */
    x_emit(x, op, 0, 0);
  }
}

static int x_compile_named(const char *src, xprog_t *x, char (*names)[72], int nnames) {
  memset(x, 0, sizeof *x);
  x->names = names;
  x->nnames = nnames;
  x->p = src;
  x_expr(x, 0);
  x_ws(x);
  if (!x->err && *x->p) x->err = "unexpected characters";
  if (!x->err && x->n == 0) x->err = "empty expression";
  if (x->err) { P_ERR("cannot parse expression `%s': %s\n", src, x->err); free(x->code); x->code = NULL; return FCFC_ERR_AST; }
  return 0;
}

/* a column value or an intermediate of SELECTION: long, double or bool */
enum { XT_LONG, XT_DBL, XT_BOOL };
typedef struct { int t; long l; double d; } xval_t;

static double xv_dbl(const xval_t *v) { return v->t == XT_DBL ? v->d : (double) v->l; }
static int xv_true(const xval_t *v) { return v->t == XT_DBL ? v->d != 0 : v->l != 0; }

/* integer power with libast's conventions (libast.c:ipow): exponents
 * above 255 give 1, +-1 or 0, the exponent is taken modulo 256, and from
 * 64 on only bases 1 and -1 survive; products wrap around */
static long x_ipow(long b, long e) {
  if (e > 255) return b == 1 ? 1 : (b == -1 ? 1 - 2 * (e & 1) : 0);
  unsigned ue = (unsigned char) e;
  if (ue >= 64) return b == 1 ? 1 : (b == -1 ? 1 - 2 * (long) (ue & 1) : 0);
  unsigned long r = 1, x = (unsigned long) b;
  for (; ue; ue >>= 1, x *= x) if (ue & 1) r *= x;
  return (long) r;
}

/* SELECTION with libast's dynamic types (libast.c:ast_eval_bool) */
static int x_eval_sel(const xprog_t *x, const xval_t *col, xval_t *st) {
  int sp = 0;
  for (int i = 0; i < x->n; i++) {
    const xop_t *o = &x->code[i];
    xval_t *v;
    switch (o->op) {
      case X_NUM:
        st[sp].t = o->islong ? XT_LONG : XT_DBL; st[sp].l = o->lv; st[sp].d = o->v; sp++;
        break;
      case X_COL: st[sp++] = col[o->col - 1]; break;
      case X_NEG: case X_ABS: case X_SQRT: case X_LN: case X_LOG: case X_NOT: case X_ISFIN:
        v = &st[sp - 1];
        if (o->op == X_NOT) { v->l = !xv_true(v); v->t = XT_BOOL; }
        else if (o->op == X_ISFIN) { v->l = v->t == XT_DBL ? isfinite(v->d) != 0 : 1; v->t = XT_BOOL; }
        else if ((o->op == X_NEG || o->op == X_ABS) && v->t == XT_LONG)
          v->l = (o->op == X_NEG || v->l < 0) ? -v->l : v->l;
        else {
          const double a = xv_dbl(v);
          v->d = (o->op == X_NEG) ? -a : (o->op == X_ABS) ? fabs(a) : (o->op == X_SQRT) ? sqrt(a)
                 : (o->op == X_LN) ? log(a) : log10(a);
          v->t = XT_DBL;
        }
        break;
      default: {
        xval_t *a = &st[sp - 2], *b = &st[sp - 1];
        sp--;
        if (o->op == X_AND || o->op == X_OR) {
          const int r = (o->op == X_AND) ? (xv_true(a) && xv_true(b)) : (xv_true(a) || xv_true(b));
          a->t = XT_BOOL; a->l = r;
          break;
        }
        const int lng = (a->t != XT_DBL && b->t != XT_DBL);
        if (o->op >= X_LT) {
          int r;
          if (lng) {
            switch (o->op) {
              case X_LT: r = a->l < b->l; break;
              case X_LE: r = a->l <= b->l; break;
              case X_GT: r = a->l > b->l; break;
              case X_GE: r = a->l >= b->l; break;
              case X_EQ: r = a->l == b->l; break;
              default: r = a->l != b->l; break;
            }
          }
          else {
            const double u = xv_dbl(a), w = xv_dbl(b);
            switch (o->op) {
              case X_LT: r = u < w; break;
              case X_LE: r = u <= w; break;
              case X_GT: r = u > w; break;
              case X_GE: r = u >= w; break;
              case X_EQ: r = u == w; break;
              default: r = u != w; break;
            }
          }
          a->t = XT_BOOL; a->l = r;
          break;
        }
        if (lng) {
          const long u = a->l, w = b->l;
          if ((o->op == X_DIV || o->op == X_REM) && w == 0) return -1;
          switch (o->op) {
            case X_POW: a->l = x_ipow(u, w); break;
            case X_MUL: a->l = (long) ((unsigned long) u * (unsigned long) w); break;
            case X_DIV: a->l = u / w; break;
            case X_REM: a->l = u % w; break;
            case X_ADD: a->l = (long) ((unsigned long) u + (unsigned long) w); break;
            default: a->l = (long) ((unsigned long) u - (unsigned long) w); break;
          }
          a->t = XT_LONG;
        }
        else {
          const double u = xv_dbl(a), w = xv_dbl(b);
          switch (o->op) {
            case X_POW: a->d = pow(u, w); break;
            case X_MUL: a->d = u * w; break;
            case X_DIV: a->d = u / w; break;
            case X_REM: a->d = fmod(u, w); break;
            case X_ADD: a->d = u + w; break;
            default: a->d = u - w; break;
          }
          a->t = XT_DBL;
        }
      }
    }
  }
  return xv_true(&st[0]);
}

static double x_eval(const xprog_t *x, const double *col, double *st) {
  int sp = 0;
  for (int i = 0; i < x->n; i++) {
    const xop_t *o = &x->code[i];
    double a, b;
    switch (o->op) {
      case X_NUM: st[sp++] = o->v; break;
      case X_COL: st[sp++] = col[o->col - 1]; break;
      case X_NEG: st[sp - 1] = -st[sp - 1]; break;
      case X_NOT: st[sp - 1] = !st[sp - 1]; break;
      case X_ABS: st[sp - 1] = fabs(st[sp - 1]); break;
      case X_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
      case X_LN: st[sp - 1] = log(st[sp - 1]); break;
      case X_LOG: st[sp - 1] = log10(st[sp - 1]); break;
      case X_ISFIN: st[sp - 1] = isfinite(st[sp - 1]) ? 1 : 0; break;
      default:
        b = st[--sp]; a = st[sp - 1];
        switch (o->op) {
          case X_POW: a = pow(a, b); break;
          case X_MUL: a = a * b; break;
          case X_DIV: a = a / b; break;
          case X_REM: a = fmod(a, b); break;
          case X_ADD: a = a + b; break;
          case X_SUB: a = a - b; break;
          case X_LT: a = a < b; break;
          case X_LE: a = a <= b; break;
          case X_GT: a = a > b; break;
          case X_GE: a = a >= b; break;
          case X_EQ: a = a == b; break;
          case X_NE: a = a != b; break;
          case X_AND: a = (a != 0) && (b != 0); break;
          default: a = (a != 0) || (b != 0); break;
        }
        st[sp - 1] = a;
    }
  }
  return st[0];
}

/* The conversions of an ASCII_FORMATTER, split as ascii_fmtr.c splits
 * them: each piece is the literal text before one conversion plus the
 * conversion, scanned with sscanf and "%n" for its length.  Its type
 * decides how the column reaches the expressions (read_ascii.c:
 * ascii_col_init): %d / %i are int, %ld / %li long, %f %e %g %a float,
 * with l double; strings and the other types are columns that are read
 * past but cannot be used; %* conversions are not columns. */
enum { FT_SKIP, FT_NULL, FT_STR, FT_INT, FT_LONG, FT_FLT, FT_DBL };
/* fast: the piece is only blanks and one %d / %i / %ld / %li / %f / %lf
 * (e, g, a alike) / %s conversion without width, suppressed or not, which
 * is scanned with strtol / strtof / strtod -- the conversions sscanf itself
 * applies (C99 7.19.6.2) -- instead of re-reading the format every row;
 * 'i' when the base is detected from the prefix */
typedef struct { char *fmt; int type; int fast; char conv; int base; } fseg_t;

static void fsegs_free(fseg_t *f, int n) { for (int i = 0; i < n; i++) free(f[i].fmt); free(f); }

static int fmt_split(const char *fmt, fseg_t **out) {
  int n = 0, cap = 8;
  fseg_t *f = malloc(sizeof(fseg_t) * cap);
  const char *seg = fmt, *p = fmt;
  while (*p) {
    if (*p != '%') { p++; continue; }
    if (p[1] == '%') { p += 2; continue; }
    const char *pc = p++;                      /* the '%' */
    int supp = 0;
    if (*p == '*') { supp = 1; p++; }
    while (isdigit((unsigned char) *p)) p++;
    char lm = 0;
    if (*p == 'h' || *p == 'l') { lm = *p++; if (*p == lm) { lm = (char) (lm == 'h' ? 'H' : 'L'); p++; } }
    else if (*p == 'j' || *p == 'z' || *p == 't' || *p == 'L') { lm = (*p == 'L') ? 'D' : *p; p++; }
    const char cv = *p;
    if (!cv) break;
    int type = FT_NULL;
    if (cv == '[') { p++; if (*p == '^') p++; if (*p == ']') p++; while (*p && *p != ']') p++; }
    if (cv == 'd' || cv == 'i') type = (lm == 0) ? FT_INT : (lm == 'l') ? FT_LONG : FT_NULL;
    else if (strchr("aAeEfFgG", cv)) type = (lm == 0) ? FT_FLT : (lm == 'l') ? FT_DBL : FT_NULL;
    else if (cv == 's' || cv == 'c' || cv == '[') type = (lm == 0) ? FT_STR : FT_NULL;
    if (*p) p++;
    if (supp) type = FT_SKIP;
    /* the piece, with '*' for values that are not stored, and "%n" */
    const size_t pre = (size_t) (pc - seg), rest = (size_t) (p - pc - 1);
    char *t = malloc(pre + rest + 6);
    memcpy(t, seg, pre);
    size_t k = pre;
    t[k++] = '%';
    if (!supp && (type == FT_NULL || type == FT_STR)) t[k++] = '*';
    memcpy(t + k, pc + 1, rest);
    k += rest;
    memcpy(t + k, "%n", 3);
    if (n == cap) { cap *= 2; f = realloc(f, sizeof(fseg_t) * cap); }
    f[n].fmt = t; f[n].type = type; n++;
    /* the fast path applies without a width, with blanks only before the % */
    int blanks = 1;
    for (const char *q = seg; q < pc; q++) if (!isspace((unsigned char) *q)) { blanks = 0; break; }
    const char *w = pc + 1 + supp;
    f[n - 1].fast = blanks && !isdigit((unsigned char) *w) && (lm == 0 || lm == 'l') &&
                    (strchr("dieEfFgGaA", cv) || (cv == 's' && lm == 0));
    f[n - 1].conv = (cv == 's') ? 's' : (cv == 'd' || cv == 'i') ? 'd' : 'f';
    f[n - 1].base = (cv == 'i') ? 0 : 10;
    seg = p;
  }
  *out = f;
  return n;
}

/* One row of an ASCII catalogue into the typed fields fv (the formatter's
 * pieces, or whitespace/comma fields without a formatter).  0: the row is
 * read; FCFC_ERR_ASCII: it does not match. */
static int ascii_fields(const char *line, const fseg_t *fs, int nf, int nfield, xval_t *fv) {
  if (!fs) {
    const char *q = line;
    for (int ti = 0; ti < nfield; ti++) {
      while (*q == ' ' || *q == '\t' || *q == '\r' || *q == ',') q++;
      char *t;
      fv[ti].t = XT_DBL;
      fv[ti].d = strtod(q, &t);
      if (!*q) return FCFC_ERR_ASCII;
      if (t == q) while (*t && !isspace((unsigned char) *t) && *t != ',') t++;
      q = t;
    }
    return 0;
  }
  /* one conversion at a time (read_ascii.c:ascii_read_line) */
  const char *q = line;
  for (int i = 0; i < nf; i++) {
    int n = 0, r, iv;
    long lv;
    float fv32;
    double dv;
    if (fs[i].fast) {
      const char *b = q;
      char *t;
      while (isspace((unsigned char) *b)) b++;
      if (fs[i].conv == 's') {
        t = (char *) b;
        while (*t && !isspace((unsigned char) *t)) t++;
      }
      else if (fs[i].conv == 'd') {
        lv = strtol(b, &t, fs[i].base);
        fv[i].t = XT_LONG; fv[i].l = (fs[i].type == FT_INT) ? (int) lv : lv;
      }
      else if (fs[i].type == FT_FLT) { fv32 = strtof(b, &t); fv[i].t = XT_DBL; fv[i].d = (double) fv32; }
      else { dv = strtod(b, &t); fv[i].t = XT_DBL; fv[i].d = dv; }
      if (t == b) return FCFC_ERR_ASCII;
      q = t;
      continue;
    }
    switch (fs[i].type) {
      case FT_INT: r = sscanf(q, fs[i].fmt, &iv, &n); fv[i].t = XT_LONG; fv[i].l = iv; break;
      case FT_LONG: r = sscanf(q, fs[i].fmt, &lv, &n); fv[i].t = XT_LONG; fv[i].l = lv; break;
      case FT_FLT: r = sscanf(q, fs[i].fmt, &fv32, &n); fv[i].t = XT_DBL; fv[i].d = (double) fv32; break;
      case FT_DBL: r = sscanf(q, fs[i].fmt, &dv, &n); fv[i].t = XT_DBL; fv[i].d = dv; break;
      case FT_STR:
        while (isspace((unsigned char) *q)) q++;
        r = sscanf(q, fs[i].fmt, &n) + 1;
        break;
      default: r = sscanf(q, fs[i].fmt, &n) + 1; break;
    }
    if (r != 1 || n == 0) return FCFC_ERR_ASCII;
    q += n;
  }
  return 0;
}

/* the rows one thread reads: growable coordinate arrays */
typedef struct { double *v[4]; int64_t n, cap; long bad; int err; } rows_t;

static void rows_push(rows_t *r, int nout, const double *v) {
  if (r->n == r->cap) {
    r->cap = r->cap ? 2 * r->cap : 4096;
    for (int k = 0; k < nout; k++) r->v[k] = realloc(r->v[k], sizeof(double) * (size_t) r->cap);
  }
  for (int k = 0; k < nout; k++) r->v[k][r->n] = v[k];
  r->n++;
}

/* Read an ASCII catalogue: nout expressions (x, y, z[, w]) per selected
 * row (read_ascii.c:read_ascii_data).  Columns are the formatter's
 * conversions that are not suppressed ($n: the n-th of them); without a
 * formatter, whitespace- or comma-separated fields read as double.  The
 * file is read in blocks of whole lines, each block's lines split among
 * the OpenMP threads and their rows appended in file order. */
static int read_ascii(const char *fname, long skip, char cmt, const char *fmt,
    xprog_t *out, int nout, const xprog_t *sel, cat_t *c) {
  FILE *fp = fopen(fname, "rb");
  if (!fp) { P_ERR("cannot open catalogue `%s'\n", fname); return FCFC_ERR_FILE; }
  fseg_t *fs = NULL;
  int nf = fmt ? fmt_split(fmt, &fs) : 0;
  int maxcol = sel ? sel->maxcol : 0, depth = sel ? sel->maxdepth : 1;
  for (int k = 0; k < nout; k++) {
    if (out[k].maxcol > maxcol) maxcol = out[k].maxcol;
    if (out[k].maxdepth > depth) depth = out[k].maxdepth;
  }
  /* the formatter piece (or the field) of column n (1-based); only the
   * columns the expressions use must be numeric */
  int *field = malloc(sizeof(int) * (maxcol + 1));
  char *used = calloc((size_t) maxcol + 1, 1);
  for (int k = 0; k <= nout; k++) {
    const xprog_t *xp = (k < nout) ? &out[k] : sel;
    if (xp) for (int i = 0; i < xp->n; i++) if (xp->code[i].op == X_COL) used[xp->code[i].col] = 1;
  }
  int nfield = 0;
  for (int cn = 1; cn <= maxcol; cn++) {
    if (!fmt) { field[cn] = cn - 1; continue; }
    int seen = 0; field[cn] = -1;
    for (int i = 0; i < nf; i++) {
      if (fs[i].type == FT_SKIP) continue;
      if (++seen == cn) {
        if (used[cn] && (fs[i].type == FT_STR || fs[i].type == FT_NULL)) {
          P_ERR("column $%d of ASCII_FORMATTER is not an int, long, float or double conversion\n", cn);
          free(field); free(used); fsegs_free(fs, nf); fclose(fp); return FCFC_ERR_CFG;
        }
        field[cn] = i; break;
      }
    }
    if (field[cn] < 0) {
      P_ERR("column $%d not in ASCII_FORMATTER\n", cn);
      free(field); free(used); fsegs_free(fs, nf); fclose(fp); return FCFC_ERR_CFG;
    }
  }
  if (!fmt) for (int cn = 1; cn <= maxcol; cn++) if (field[cn] + 1 > nfield) nfield = field[cn] + 1;
  if (fmt) nfield = nf;

  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  rows_t *part = calloc((size_t) nth, sizeof(rows_t));
  c->x = c->y = c->z = c->w = NULL;
  c->n = 0;
  int64_t cap = 0;
  const size_t BLOCK = (size_t) 32 << 20;
  char *buf = malloc(BLOCK + 1);
  size_t have = 0;
  long lines_before = 0;           /* lines of the file before this block */
  int e = 0, eof = 0;
  /* ASCII_SKIP: the first lines are not read */
  for (int ch; lines_before < skip && (ch = fgetc(fp)) != EOF;) if (ch == '\n') lines_before++;
  while (!e && !eof) {
    const size_t want = BLOCK - have;
    const size_t got = fread(buf + have, 1, want, fp);
    have += got;
    eof = got < want;
    /* whole lines: up to the last newline (all of it at the end of file) */
    size_t len = have;
    if (!eof) {
      while (len > 0 && buf[len - 1] != '\n') len--;
      if (len == 0) { P_ERR("line too long in `%s'\n", fname); e = FCFC_ERR_ASCII; break; }
    }
    if (len == 0) break;
    const char saved = buf[len];
    buf[len] = '\0';
    /* split the block among the threads at line starts */
    size_t *cut = malloc(sizeof(size_t) * (size_t) (nth + 1));
    long *nlines = calloc((size_t) nth, sizeof(long));
    cut[0] = 0;
    for (int t = 1; t < nth; t++) {
      size_t p = len * (size_t) t / (size_t) nth;
      if (p < cut[t - 1]) p = cut[t - 1];
      while (p < len && p > 0 && buf[p - 1] != '\n') p++;
      cut[t] = p;
    }
    cut[nth] = len;
    for (int t = 0; t < nth; t++) { part[t].n = 0; part[t].err = 0; part[t].bad = 0; }
#ifdef _OPENMP
#pragma omp parallel for schedule(static, 1) num_threads(nth)
#endif
    for (int t = 0; t < nth; t++) {
      xval_t *fv = malloc(sizeof(xval_t) * (size_t) (nfield + 1));
      xval_t *col = malloc(sizeof(xval_t) * (size_t) (maxcol + 1));
      xval_t *xst = malloc(sizeof(xval_t) * (size_t) (depth + 1));
      double *cold = malloc(sizeof(double) * (size_t) (maxcol + 1));
      double *stack = malloc(sizeof(double) * (size_t) (depth + 1));
      long lno = 0;                    /* lines of this part so far */
      for (size_t p = cut[t]; p < cut[t + 1] && !part[t].err;) {
        char *line = buf + p;
        char *nl = memchr(line, '\n', cut[t + 1] - p);
        size_t next = nl ? (size_t) (nl - buf) + 1 : cut[t + 1];
        if (nl) *nl = '\0';
        p = next;
        lno++;
        const char *q = line;
        while (isspace((unsigned char) *q)) q++;
        if (!*q || (cmt && *q == cmt)) continue;
        if (ascii_fields(line, fmt ? fs : NULL, nf, nfield, fv)) { part[t].err = FCFC_ERR_ASCII; part[t].bad = lno; break; }
        for (int cn = 1; cn <= maxcol; cn++) {
          if (!used[cn]) continue;
          col[cn - 1] = fv[field[cn]];
          cold[cn - 1] = xv_dbl(&col[cn - 1]);
        }
        if (sel) {
          const int r = x_eval_sel(sel, col, xst);
          if (r < 0) { part[t].err = FCFC_ERR_AST; part[t].bad = lno; break; }
          if (!r) continue;
        }
        double v[4];
        for (int k = 0; k < nout; k++) v[k] = x_eval(&out[k], cold, stack);
        rows_push(&part[t], nout, v);
      }
      free(fv); free(col); free(xst); free(cold); free(stack);
      nlines[t] = lno;
    }
    for (int t = 0; t < nth && !e; t++) {
      if (part[t].err) {
        part[t].bad += lines_before;
        for (int u = 0; u < t; u++) part[t].bad += nlines[u];
        if (part[t].err == FCFC_ERR_AST)
          P_ERR("integer division by zero in SELECTION (line %ld of `%s')\n", part[t].bad, fname);
        else P_ERR("line %ld of `%s' does not match ASCII_FORMATTER\n", part[t].bad, fname);
        e = part[t].err;
        break;
      }
      if (c->n + part[t].n > cap) {
        while (c->n + part[t].n > cap) cap = cap ? 2 * cap : 65536;
        c->x = realloc(c->x, sizeof(double) * (size_t) cap);
        c->y = realloc(c->y, sizeof(double) * (size_t) cap);
        c->z = realloc(c->z, sizeof(double) * (size_t) cap);
        if (nout > 3) c->w = realloc(c->w, sizeof(double) * (size_t) cap);
      }
      double *dst[4] = {c->x, c->y, c->z, c->w};
      for (int k = 0; k < nout; k++) memcpy(dst[k] + c->n, part[t].v[k], sizeof(double) * (size_t) part[t].n);
      c->n += part[t].n;
    }
    for (int t = 0; t < nth; t++) lines_before += nlines[t];
    free(cut); free(nlines);
    buf[len] = saved;
    memmove(buf, buf + len, have - len);
    have -= len;
  }
  if (!c->x) {
    c->x = malloc(sizeof(double)); c->y = malloc(sizeof(double)); c->z = malloc(sizeof(double));
    if (nout > 3) c->w = malloc(sizeof(double));
  }
  for (int t = 0; t < nth; t++) for (int k = 0; k < 4; k++) free(part[t].v[k]);
  free(part); free(buf); free(field); free(used);
  if (fs) fsegs_free(fs, nf);
  fclose(fp);
  if (e) return e;
  c->loaded = 1;
  return 0;
}

/* FITS binary tables (CATALOG_TYPE 1; the reference reads them with
 * cfitsio, src/io/read_fits.c).  Written from the FITS standard: header
 * blocks of 2880 bytes holding 80-character cards, the first HDU with data
 * must be a BINTABLE extension, rows of NAXIS1 bytes, big-endian fields.
 * Columns are TTYPEn (${NAME}, without regard to case) or numbers;
 * B, I, J and K columns are integers (long in SELECTION), E float, D
 * double -- the types read_fits.c:fits_read_real accepts; other types,
 * vector columns and scaled columns (TSCALn, TZEROn) are refused when an
 * expression uses them. */
typedef struct { char name[72]; char type; long repeat; long off; int scaled; } fcol_t;
typedef struct { FILE *fp; long nrow, rowlen, data; int ncol; char (*names)[72]; fcol_t *col; } fits_t;

static void fits_close(fits_t *f) {
  if (f->fp) fclose(f->fp);
  free(f->col); free(f->names);
  memset(f, 0, sizeof *f);
}

/* value of a header card: the text after "= " up to a comment, unquoted */
static void fits_value(const char *card, char *val) {
  const char *p = card + 10, *end = card + 80;
  while (p < end && *p == ' ') p++;
  size_t k = 0;
  if (p < end && *p == '\'') {
    for (p++; p < end; p++) {
      if (*p == '\'') { if (p + 1 < end && p[1] == '\'') { val[k++] = '\''; p++; continue; } break; }
      val[k++] = *p;
    }
    while (k && val[k - 1] == ' ') k--;
  }
  else {
    while (p < end && *p != '/' && *p != ' ') val[k++] = *p++;
  }
  val[k] = '\0';
}

static int fits_open(const char *fname, fits_t *f) {
  memset(f, 0, sizeof *f);
  if (strchr(fname, '[')) { P_ERR("FITS extended file names are not supported: `%s'\n", fname); return FCFC_ERR_FILE; }
  if (!(f->fp = fopen(fname, "rb"))) { P_ERR("cannot open catalogue `%s'\n", fname); return FCFC_ERR_FILE; }
  char block[2880], val[80];
  long pos = 0;
  for (int hdu = 0;; hdu++) {
    long bitpix = 0, naxis = 0, nax[9] = {0}, pcount = 0, gcount = 1, tfields = 0;
    int bintable = 0, end = 0;
    fcol_t *col = NULL;
    while (!end) {
      if (fread(block, 1, 2880, f->fp) != 2880) {
        free(col);
        P_ERR("no binary table in FITS file `%s'\n", fname);
        fits_close(f); return FCFC_ERR_FILE;
      }
      pos += 2880;
      for (int i = 0; i < 36 && !end; i++) {
        const char *card = block + 80 * i;
        char key[9];
        memcpy(key, card, 8); key[8] = '\0';
        for (int k = 7; k >= 0 && key[k] == ' '; k--) key[k] = '\0';
        if (!strcmp(key, "END")) { end = 1; break; }
        if (card[8] != '=' || card[9] != ' ') continue;
        fits_value(card, val);
        if (hdu == 0 && !strcmp(key, "SIMPLE") && strcmp(val, "T")) {
          P_ERR("`%s' is not a FITS file\n", fname); free(col); fits_close(f); return FCFC_ERR_FILE;
        }
        if (!strcmp(key, "XTENSION")) bintable = !strcmp(val, "BINTABLE");
        else if (!strcmp(key, "BITPIX")) bitpix = atol(val);
        else if (!strcmp(key, "NAXIS")) naxis = atol(val);
        else if (!strncmp(key, "NAXIS", 5) && isdigit((unsigned char) key[5])) {
          int a = atoi(key + 5);
          if (a >= 1 && a <= 8) nax[a] = atol(val);
        }
        else if (!strcmp(key, "PCOUNT")) pcount = atol(val);
        else if (!strcmp(key, "GCOUNT")) gcount = atol(val);
        else if (!strcmp(key, "TFIELDS")) {
          tfields = atol(val);
          if (tfields < 0 || tfields > 999) { P_ERR("bad TFIELDS in `%s'\n", fname); free(col); fits_close(f); return FCFC_ERR_FILE; }
          col = calloc((size_t) tfields + 1, sizeof(fcol_t));
          for (long c = 0; c < tfields; c++) col[c].repeat = -1;
        }
        else if (col && (!strncmp(key, "TTYPE", 5) || !strncmp(key, "TFORM", 5) ||
                         !strncmp(key, "TSCAL", 5) || !strncmp(key, "TZERO", 5))) {
          long c = atol(key + 5) - 1;
          if (c < 0 || c >= tfields) continue;
          if (key[1] == 'T') snprintf(col[c].name, sizeof col[c].name, "%.71s", val);
          else if (key[1] == 'F') {
            char *q;
            long r = strtol(val, &q, 10);
            col[c].repeat = (q == val) ? 1 : r;
            col[c].type = *q;
          }
          else if ((key[1] == 'S' && atof(val) != 1.0) || (key[1] == 'Z' && atof(val) != 0.0))
            col[c].scaled = 1;
        }
      }
    }
    long size = 0;
    if (naxis > 0) {
      size = 1;
      for (long a = 1; a <= naxis && a <= 8; a++) size *= nax[a];
      size = (labs(bitpix) / 8) * gcount * (pcount + size);
    }
    if (bintable) {
      f->nrow = nax[2]; f->rowlen = nax[1]; f->ncol = (int) tfields; f->col = col; f->data = pos;
      long off = 0;
      f->names = calloc((size_t) tfields + 1, sizeof *f->names);
      for (long c = 0; c < tfields; c++) {
        static const char tc[] = "LXBIJKAEDCMPQ";
        static const int tw[] = {1, 0, 1, 2, 4, 8, 1, 4, 8, 8, 16, 8, 16};
        const char *t = col[c].type ? strchr(tc, col[c].type) : NULL;
        if (!t || col[c].repeat < 0) { P_ERR("bad TFORM%ld in `%s'\n", c + 1, fname); fits_close(f); return FCFC_ERR_FILE; }
        col[c].off = off;
        off += (col[c].type == 'X') ? (col[c].repeat + 7) / 8 : col[c].repeat * tw[t - tc];
        memcpy(f->names[c], col[c].name, sizeof col[c].name);
      }
      if (off != f->rowlen) { P_ERR("row length of `%s' does not match its columns\n", fname); fits_close(f); return FCFC_ERR_FILE; }
      return 0;
    }
    free(col);
    if (hdu == 0 && size > 0) {
      P_ERR("the first HDU with data in `%s' is not a binary table\n", fname);
      fits_close(f); return FCFC_ERR_FILE;
    }
    /* skip the data of this HDU */
    const long skip = (size + 2879) / 2880 * 2880;
    if (fseek(f->fp, skip, SEEK_CUR)) { P_ERR("truncated FITS file `%s'\n", fname); fits_close(f); return FCFC_ERR_FILE; }
    pos += skip;
  }
}

static uint64_t fits_be(const unsigned char *b, int n) {
  uint64_t v = 0;
  for (int i = 0; i < n; i++) v = (v << 8) | b[i];
  return v;
}

/* the rows of a FITS table: nout expressions per selected row */
static int read_fits(const char *fname, fits_t *f, xprog_t *out, int nout, const xprog_t *sel,
    cat_t *c) {
  int maxcol = sel ? sel->maxcol : 0, depth = sel ? sel->maxdepth : 1;
  for (int k = 0; k < nout; k++) {
    if (out[k].maxcol > maxcol) maxcol = out[k].maxcol;
    if (out[k].maxdepth > depth) depth = out[k].maxdepth;
  }
  if (maxcol > f->ncol) { P_ERR("column %d not in FITS table `%s'\n", maxcol, fname); return FCFC_ERR_CFG; }
  char *used = calloc((size_t) maxcol + 1, 1);
  for (int k = 0; k <= nout; k++) {
    const xprog_t *xp = (k < nout) ? &out[k] : sel;
    if (xp) for (int i = 0; i < xp->n; i++) if (xp->code[i].op == X_COL) used[xp->code[i].col] = 1;
  }
  for (int cn = 1; cn <= maxcol; cn++) {
    if (!used[cn]) continue;
    const fcol_t *fc = &f->col[cn - 1];
    if (!strchr("BIJKED", fc->type) || fc->repeat != 1 || fc->scaled) {
      P_ERR("FITS column `%s' of `%s' is not a scalar B, I, J, K, E or D column without scaling\n",
            fc->name, fname);
      free(used); return FCFC_ERR_CFG;
    }
  }
  xval_t *col = malloc(sizeof(xval_t) * (maxcol + 1)), *xst = malloc(sizeof(xval_t) * (depth + 1));
  double *cold = malloc(sizeof(double) * (maxcol + 1)), *stack = malloc(sizeof(double) * (depth + 1));
  const long chunk = 4096;
  unsigned char *buf = malloc((size_t) (f->rowlen ? f->rowlen : 1) * chunk);
  size_t cap = (size_t) (f->nrow > 0 ? f->nrow : 1);
  c->x = malloc(sizeof(double) * cap); c->y = malloc(sizeof(double) * cap);
  c->z = malloc(sizeof(double) * cap); c->w = (nout > 3) ? malloc(sizeof(double) * cap) : NULL;
  c->n = 0;
  int e = 0;
  if (fseek(f->fp, f->data, SEEK_SET)) e = FCFC_ERR_FILE;
  for (long r0 = 0; r0 < f->nrow && !e; r0 += chunk) {
    const long nr = (f->nrow - r0 < chunk) ? f->nrow - r0 : chunk;
    if (fread(buf, (size_t) f->rowlen, (size_t) nr, f->fp) != (size_t) nr) {
      P_ERR("truncated FITS table `%s'\n", fname); e = FCFC_ERR_FILE; break;
    }
    for (long r = 0; r < nr && !e; r++) {
      const unsigned char *row = buf + r * f->rowlen;
      for (int cn = 1; cn <= maxcol; cn++) {
        if (!used[cn]) continue;
        const fcol_t *fc = &f->col[cn - 1];
        const unsigned char *b = row + fc->off;
        xval_t *v = &col[cn - 1];
        uint64_t u;
        switch (fc->type) {
          case 'B': v->t = XT_LONG; v->l = b[0]; break;
          case 'I': v->t = XT_LONG; v->l = (int16_t) fits_be(b, 2); break;
          case 'J': v->t = XT_LONG; v->l = (int32_t) fits_be(b, 4); break;
          case 'K': v->t = XT_LONG; v->l = (long) (int64_t) fits_be(b, 8); break;
          case 'E': { uint32_t w32 = (uint32_t) fits_be(b, 4); float fv; memcpy(&fv, &w32, 4); v->t = XT_DBL; v->d = fv; break; }
          default: { u = fits_be(b, 8); double dv; memcpy(&dv, &u, 8); v->t = XT_DBL; v->d = dv; break; }
        }
        cold[cn - 1] = xv_dbl(v);
      }
      if (sel) {
        const int rs = x_eval_sel(sel, col, xst);
        if (rs < 0) { P_ERR("integer division by zero in SELECTION (row %ld of `%s')\n", r0 + r + 1, fname); e = FCFC_ERR_AST; break; }
        if (!rs) continue;
      }
      c->x[c->n] = x_eval(&out[0], cold, stack);
      c->y[c->n] = x_eval(&out[1], cold, stack);
      c->z[c->n] = x_eval(&out[2], cold, stack);
      if (c->w) c->w[c->n] = x_eval(&out[3], cold, stack);
      c->n++;
    }
  }
  free(used); free(col); free(xst); free(cold); free(stack); free(buf);
  if (e) return e;
  c->loaded = 1;
  return 0;
}

#if FCFC_CLI_GEOM == 1
/* the first two columns of an ASCII file (Z_CMVDST_FILE): lines that are
 * empty or start with '#' are skipped (read_ascii.c:read_ascii_table) */
static int read_table2(const char *fname, double **x, double **y, int64_t *n) {
  FILE *fp = fopen(fname, "r");
  if (!fp) { P_ERR("cannot open file for reading: `%s'\n", fname); return FCFC_ERR_FILE; }
  size_t cap = 1024, m = 0;
  double *a = malloc(sizeof(double) * cap), *b = malloc(sizeof(double) * cap);
  char *line = NULL; size_t lcap = 0;
  int e = 0;
  while (getline(&line, &lcap, fp) >= 0) {
    char *p = trim(line);
    if (!*p || *p == '#') continue;
    if (m == cap) { cap *= 2; a = realloc(a, sizeof(double) * cap); b = realloc(b, sizeof(double) * cap); }
    if (sscanf(p, "%lf %lf", a + m, b + m) != 2) { P_ERR("failed to read line: %s\n", p); e = FCFC_ERR_FILE; break; }
    m++;
  }
  free(line);
  fclose(fp);
  *x = a; *y = b; *n = (int64_t) m;
  return e;
}
#endif

/* ---------------------------------------------------------------------- */
/* pair-count files: reader for the resume path (OVERWRITE = 1).           */
/* As read_pair_count (src/fcfc/2pt_box/read_res.c:256-560, survey twin    */
/* src/fcfc/2pt/read_res.c) it takes the NORMALISED counts from the file   */
/* and checks the settings against the current configuration: the header   */
/* line "# Created by <code> <version> (<vernum>) ; format = <0|1>", then  */
/*  binary: blocks [u64 size][payload][u64 size] read by position --       */
/*    spec int32[7] (iso) or [8]; precision spec[0] (0: FP32 edges/box);    */
/*    spec[2] periodic (box) or not (survey); spec[5] binning; spec[6] ns;  */
/*    spec[7] nmu or np -- the box size (rescaled, box only), the s edges,  */
/*    the pi edges (s_perp,pi), the object numbers (1 or 2, u64 or f64),    */
/*    the normalisation, the normalised counts, the raw counts; edges and   */
/*    box within 1e-10 (1e-5 for FP32 files);                              */
/*  ASCII: comment lines with the column header on the 4th line, then ntot  */
/*    rows whose last column is the normalised count.                      */
/* Unlike the reference, a file that ends before its count blocks, or a    */
/* pi-bin number that differs (reported there without stopping), fails.   */
static int read_pair_file(const char *fname, const fcfc_bins *b, const double *bres,
    double *ncnt) {
  FILE *fp = fopen(fname, "rb");
  if (!fp) { P_ERR("cannot open the pair count file `%s'\n", fname); return FCFC_ERR_FILE; }
  char head[1024];
  if (!fgets(head, sizeof head, fp)) { fclose(fp); P_ERR("empty pair count file `%s'\n", fname); return FCFC_ERR_FILE; }
  char code[64];
  int ver = 0, fmt = -1;
  if (sscanf(head, "# Created by %63s %*s (%d) ; format = %d", code, &ver, &fmt) != 3 ||
      strcmp(code, CODE_NAME)) {
    fclose(fp); P_ERR("unrecognized pair count file: `%s'\n", fname); return FCFC_ERR_FILE;
  }
  const int spi = (b->bintype == FCFC_BIN_SPI), iso = (b->bintype == FCFC_BIN_ISO);
  int e = 0;
  if (fmt == 1) {                                   /* ASCII (save_res.c:45-164) */
    const char *hdr = iso ? "Columns: s_min(1) s_max(2) normalized_pair_count(3)"
        : (spi ? "Columns: s_perp_min(1) s_perp_max(2) pi_min(3) pi_max(4) normalized_pair_count(5)"
               : "Columns: s_min(1) s_max(2) mu_min(3) mu_max(4) normalized_pair_count(5)");
    char *line = NULL; size_t cap = 0;
    int lno = 1, n = 0;
    while (getline(&line, &cap, fp) >= 0) {
      lno++;
      if (line[0] == '#') {
        if (lno == 4 && strncmp(line + 2, hdr, strlen(hdr))) {
          P_ERR("unexpected separation bins in pair count file: `%s'\n", fname);
          e = FCFC_ERR_FILE; break;
        }
        continue;
      }
      if (n == 0 && lno != 5) {
        P_ERR("invalid number of header lines in pair count file: `%s'\n", fname);
        e = FCFC_ERR_FILE; break;
      }
      char *p = trim(line);
      if (!*p) continue;
      if (n == b->ntot) { P_ERR("too many rows in pair count file: `%s'\n", fname); e = FCFC_ERR_FILE; break; }
      double v;
      if (sscanf(p, iso ? "%*f %*f %lf" : "%*f %*f %*f %*f %lf", &v) != 1) {
        P_ERR("failed to read line %d of file: `%s'\n", lno, fname); e = FCFC_ERR_FILE; break;
      }
      ncnt[n++] = v;
    }
    free(line);
    fclose(fp);
    if (!e && n != b->ntot) {
      P_ERR("unexpected number of bins (%d) in pair count file: `%s'\n", n, fname);
      e = FCFC_ERR_FILE;
    }
    return e;
  }
  if (fmt != 0) { fclose(fp); P_ERR("unrecognized pair count file format: %d\n", fmt); return FCFC_ERR_FILE; }
  /* binary: block k of the layout, by position */
  const int nblk = 7 + (FCFC_CLI_GEOM == 0 ? 1 : 0) + (spi ? 1 : 0) - 1;
  int fp32 = 0, isauto = 0, wt = 0, k = 0;
  char *buf = NULL;
  uint64_t sz, sz2;
  for (; k < nblk; k++) {
    if (fread(&sz, sizeof sz, 1, fp) != 1 || sz > ((uint64_t) 1 << 32)) break;
    char *nb = realloc(buf, sz ? sz : 1);
    if (!nb) { e = FCFC_ERR_MEMORY; break; }
    buf = nb;
    if ((sz && fread(buf, sz, 1, fp) != 1) || fread(&sz2, sizeof sz2, 1, fp) != 1 || sz2 != sz) {
      P_ERR("invalid format of pair count file: `%s'\n", fname); e = FCFC_ERR_FILE; break;
    }
    /* the role of block k */
    int role = k;                          /* 0 spec, 1 box, 2 s, 3 pi, 4 n, 5 norm, 6 ncnt, 7 cnt */
    if (FCFC_CLI_GEOM != 0 && role >= 1) role++;
    if (!spi && role >= 3) role++;
    const size_t fsz = fp32 ? sizeof(float) : sizeof(double);
    if (role == 0) {
      int32_t spec[8] = {0};
      if (sz != (uint64_t) (iso ? 7 : 8) * sizeof(int32_t)) { e = FCFC_ERR_FILE; P_ERR("invalid format of pair count file: `%s'\n", fname); break; }
      memcpy(spec, buf, sz);
      fp32 = (spec[0] == 0);
      if (fp32) fprintf(stderr, "Warning: pair count file was generated with single precision: `%s'\n", fname);
      if (spec[1] == 1 && b->bintype == FCFC_BIN_SMU)
        fprintf(stderr, "Warning: pair count file was generated with `-DWITH_MU_ONE`: `%s'\n", fname);
      if (spec[2] != (FCFC_CLI_GEOM == 0 ? 1 : 0)) {
        P_ERR("%s pair counts in file: `%s'\n", FCFC_CLI_GEOM == 0 ? "non-periodic" : "periodic", fname);
        e = FCFC_ERR_FILE; break;
      }
      isauto = (spec[3] == 0);
      wt = (spec[4] == 1);
      if (spec[5] != b->bintype) { P_ERR("unexpected binning scheme in pair count file: `%s'\n", fname); e = FCFC_ERR_FILE; break; }
      if (spec[6] != b->ns) { P_ERR("unexpected number of separation bins (%d) in pair count file: `%s'\n", spec[6], fname); e = FCFC_ERR_FILE; break; }
      if (!iso && spec[7] != (spi ? b->np : b->nmu)) {
        P_ERR("unexpected number of %s bins (%d) in pair count file: `%s'\n", spi ? "pi" : "mu", spec[7], fname);
        e = FCFC_ERR_FILE; break;
      }
    }
    else if (role == 1 || role == 2 || role == 3) {
      const int len = role == 1 ? 3 : (role == 2 ? b->ns + 1 : b->np + 1);
      const double *want = role == 1 ? bres : (role == 2 ? b->sbin_raw : b->pbin_raw);
      const double tol = fp32 ? 1e-5 : 1e-10;
      if (sz != (uint64_t) len * fsz) { P_ERR("invalid format of pair count file: `%s'\n", fname); e = FCFC_ERR_FILE; break; }
      for (int i = 0; i < len; i++) {
        double v;
        if (fp32) { float f; memcpy(&f, buf + i * sizeof f, sizeof f); v = f; }
        else memcpy(&v, buf + i * sizeof v, sizeof v);
        if (fabs(v - want[i]) > tol) {
          P_ERR("unexpected %s in pair count file: `%s'\n",
                role == 1 ? "box size" : (role == 2 ? "separation bin" : "pi bin"), fname);
          e = FCFC_ERR_FILE; break;
        }
      }
      if (e) break;
    }
    else if (role == 4) {
      if (sz != (uint64_t) 8 * (isauto ? 1 : 2)) { P_ERR("invalid format of pair count file: `%s'\n", fname); e = FCFC_ERR_FILE; break; }
    }
    else if (role == 5) {
      if (sz != 8) { P_ERR("invalid format of pair count file: `%s'\n", fname); e = FCFC_ERR_FILE; break; }
    }
    else if (role == 6) {
      if (sz != (uint64_t) b->ntot * 8) { P_ERR("invalid format of pair count file: `%s'\n", fname); e = FCFC_ERR_FILE; break; }
      memcpy(ncnt, buf, sz);
    }
    else {
      if (sz != (uint64_t) b->ntot * 8) { P_ERR("invalid format of pair count file: `%s'\n", fname); e = FCFC_ERR_FILE; break; }
    }
  }
  (void) wt;
  if (!e && k < nblk) { P_ERR("pair count file ends before its counts: `%s'\n", fname); e = FCFC_ERR_FILE; }
  if (!e && fread(&sz, 1, 1, fp) == 1) { P_ERR("invalid format of pair count file: `%s'\n", fname); e = FCFC_ERR_FILE; }
  free(buf);
  fclose(fp);
  return e;
}

/* ---------------------------------------------------------------------- */
/* CF estimator expressions: + - * / ^ ( ), numbers, pair labels, @@        */
typedef struct { const char *p; const double *val; char (*lab)[3]; int npc; double rr; int err; } ex_t;
static double ex_sum(ex_t *e);
static double ex_atom(ex_t *e) {
  while (isspace((unsigned char) *e->p)) e->p++;
  if (*e->p == '(') { e->p++; double v = ex_sum(e); while (isspace((unsigned char) *e->p)) e->p++; if (*e->p == ')') e->p++; else e->err = 1; return v; }
  if (*e->p == '-') { e->p++; return -ex_atom(e); }
  if (*e->p == '+') { e->p++; return ex_atom(e); }
  if (e->p[0] == '@' && e->p[1] == '@') { e->p += 2; return e->rr; }
  if (isupper((unsigned char) e->p[0]) && isupper((unsigned char) e->p[1])) {
    for (int i = 0; i < e->npc; i++)
      if (e->lab[i][0] == e->p[0] && e->lab[i][1] == e->p[1]) { e->p += 2; return e->val[i]; }
    e->err = 2; return 0;
  }
  char *end;
  double v = strtod(e->p, &end);
  if (end == e->p) { e->err = 1; return 0; }
  e->p = end;
  return v;
}
static double ex_pow(ex_t *e) {
  double v = ex_atom(e);
  while (isspace((unsigned char) *e->p)) e->p++;
  if (*e->p == '^') { e->p++; return pow(v, ex_pow(e)); }
  return v;
}
static double ex_prod(ex_t *e) {
  double v = ex_pow(e);
  for (;;) {
    while (isspace((unsigned char) *e->p)) e->p++;
    if (*e->p == '*') { e->p++; v *= ex_pow(e); }
    else if (*e->p == '/') { e->p++; v /= ex_pow(e); }
    else return v;
  }
}
static double ex_sum(ex_t *e) {
  double v = ex_prod(e);
  for (;;) {
    while (isspace((unsigned char) *e->p)) e->p++;
    if (*e->p == '+') { e->p++; v += ex_prod(e); }
    else if (*e->p == '-') { e->p++; v -= ex_prod(e); }
    else return v;
  }
}

static double legp(int ell, double x) {      /* src/math/legpoly.h:47-61 */
  const double x2 = x * x;
  switch (ell) {
    case 0: return 1;
    case 1: return x;
    case 2: return 1.5 * x2 - 0.5;
    case 3: return 2.5 * x * (x2 - 0.6);
    case 4: return 4.375 * x2 * x2 - 3.75 * x2 + 0.375;
    case 5: return 8.75 * x * (0.9 * x2 * x2 - x2 + 0x1.b6db6db6db6dbp-3);
    case 6: return 14.4375 * x2 * x2 * x2 - 19.6875 * x2 * x2 + 6.5625 * x2 - 0.3125;
    default: return 0;
  }
}

static int file_exists(const char *f) { struct stat st; return stat(f, &st) == 0; }

static void usage(void) {
  printf("Usage: %s_GPU [OPTION]\n"
         "Pair counts and 2-point correlation functions with the MI355X engine.\n"
         "Options and configuration keys as %s (-c CONFIG_FILE, -i CATALOG, ...);\n"
         "extra: --gpu-prec GPU_PRECISION (0: FP64 kernel, 1: FP32 + FP64 edge re-evaluation)\n",
         CODE_NAME, CODE_NAME);
}

#define FAIL(code) do { ret = (code); goto done; } while (0)

int main(int argc, char **argv) {
  /* progress lines leave as they are printed, also into a pipe: a caller
   * that captures them sees every stage that has finished */
  setvbuf(stdout, NULL, _IOLBF, 0);
  int ret = 0;
  const char *conf_file = DEFAULT_CONF;
  int conf_given = 0;
  /* command line */
  for (int i = 1; i < argc; i++) {
    const char *a = argv[i];
    if (!strcmp(a, "-h") || !strcmp(a, "--help")) { usage(); return 0; }
    if (!strcmp(a, "-V") || !strcmp(a, "--version")) { printf("%s_GPU (FCFC v1.0.1 compatible)\n", CODE_NAME); return 0; }
    const opt_t *o = NULL;
    const char *val = NULL;
    if (a[0] == '-' && a[1] == '-') {
      const char *name = a + 2;
      const char *eqp = strchr(name, '=');
      size_t nl = eqp ? (size_t) (eqp - name) : strlen(name);
      for (const opt_t *q = OPTS; q->key; q++) if (q->l && strlen(q->l) == nl && !strncmp(q->l, name, nl)) { o = q; break; }
      if (eqp) val = eqp + 1;
    }
    else if (a[0] == '-' && a[1] && !a[2]) {
      for (const opt_t *q = OPTS; q->key; q++) if (q->s == a[1]) { o = q; break; }
    }
    if (!o) { P_ERR("unknown option `%s'\n", a); return FCFC_ERR_ARG; }
    if (!val) {
      if (!strcmp(o->key, "VERBOSE") && (i + 1 >= argc || argv[i + 1][0] == '-')) val = "T";
      else if (!strcmp(o->key, "PROJECTED_CF") && (i + 1 >= argc || argv[i + 1][0] == '-')) val = "T";
      else if (i + 1 < argc) val = argv[++i];
      else { P_ERR("missing value for `%s'\n", a); return FCFC_ERR_ARG; }
    }
    if (!strcmp(o->key, "CONFIG_FILE")) { conf_file = val; conf_given = 1; }
    else kv_set(o->key, val, 5);
  }
  if (read_conf_file(conf_file, conf_given)) return FCFC_ERR_FILE;
  const int verbose = get_bool("VERBOSE", 1);

  char **cats = NULL, **labels = NULL, **pos = NULL, **wts = NULL, **fmts = NULL, **pcs = NULL,
       **pcout = NULL, **cfs = NULL, **cfout = NULL, **mpout = NULL, **wpout = NULL, **poles = NULL,
       **skips = NULL, **cmts = NULL, **types = NULL;
  int ncat = 0, nlab = 0, npos = 0, nwt = 0, nfmt = 0, npc = 0, npco = 0, ncf = 0, ncfo = 0,
      nmpo = 0, nwpo = 0, npole = 0, nskip = 0, ncmt = 0, ntype = 0;
  fcfc_gpu_ctx *ctx = NULL;
  fcfc_bins bins;
  memset(&bins, 0, sizeof bins);
  cat_t *cat = NULL;
  double **cnt = NULL, **ncnt = NULL, *norm = NULL;
  int *weighted = NULL, *comp_pc = NULL, *cnvt = NULL, any_cnvt = 0, *has_wt = NULL;
  char (*pclab)[3] = NULL;
  int *pci[2] = {NULL, NULL};
  char **sels = NULL;
  int nsel = 0;
  const char *zfile = NULL;
  fcfc_cosmology cosmo;
  memset(&cosmo, 0, sizeof cosmo);

  ncat = split_list(kv_get("CATALOG"), &cats);
  if (ncat < 1) { P_ERR("CATALOG is not set\n"); return FCFC_ERR_CFG; }
  /* labels are single letters A-Z (load_conf.c:665-690) */
  if (ncat > 26) { P_ERR("there can be at most 26 catalogs set via CATALOG\n"); free_list(cats, ncat); return FCFC_ERR_CFG; }
  nlab = split_list(kv_get("CATALOG_LABEL"), &labels);
  npos = split_list(kv_get("POSITION"), &pos);
  nwt = split_list(kv_get("WEIGHT"), &wts);
  nfmt = split_list(kv_get("ASCII_FORMATTER"), &fmts);
  nskip = split_list(kv_get("ASCII_SKIP"), &skips);
  ncmt = split_list(kv_get("ASCII_COMMENT"), &cmts);
  ntype = split_list(kv_get("CATALOG_TYPE"), &types);
  npc = split_list(kv_get("PAIR_COUNT"), &pcs);
  npco = split_list(kv_get("PAIR_COUNT_FILE"), &pcout);
  ncf = split_list(kv_get("CF_ESTIMATOR"), &cfs);
  ncfo = split_list(kv_get("CF_OUTPUT_FILE"), &cfout);
  npole = split_list(kv_get("MULTIPOLE"), &poles);
  nmpo = split_list(kv_get("MULTIPOLE_FILE"), &mpout);
  nwpo = split_list(kv_get("PROJECTED_FILE"), &wpout);
  if (npos != 3 * ncat) { P_ERR("POSITION needs 3 entries per catalogue\n"); FAIL(FCFC_ERR_CFG); }
  /* CATALOG_TYPE: 0 ASCII, 1 FITS (2, HDF5, is not supported); one for
   * every catalogue or one for all */
  if (ntype > 1 && ntype != ncat) { P_ERR("CATALOG_TYPE needs one entry per catalogue\n"); FAIL(FCFC_ERR_CFG); }
  int any_ascii = (ntype == 0);
  for (int i = 0; i < ntype; i++) {
    const int t = atoi(types[i]);
    if (t == 2) { P_ERR("HDF5 catalogues (CATALOG_TYPE 2) are not supported\n"); FAIL(FCFC_ERR_CFG); }
    if (t != 0 && t != 1) { P_ERR("unknown CATALOG_TYPE %d\n", t); FAIL(FCFC_ERR_CFG); }
    if (t == 0) any_ascii = 1;
  }
  if (any_ascii && nfmt != ncat) { P_ERR("ASCII_FORMATTER is not set for every catalogue\n"); FAIL(FCFC_ERR_CFG); }
  nsel = split_list(kv_get("SELECTION"), &sels);
  /* COORD_CONVERT (survey): one flag for every catalogue or one each, with
   * the fiducial cosmology or a (redshift, distance) table
   * (2pt/load_conf.c:886-935) */
#if FCFC_CLI_GEOM == 1
  {
    char **cv; int ncv = split_list(kv_get("COORD_CONVERT"), &cv);
    cnvt = calloc(ncat, sizeof(int));
    if (ncv > 1 && ncv < ncat) { P_ERR("too few elements of COORD_CONVERT\n"); free_list(cv, ncv); FAIL(FCFC_ERR_CFG); }
    for (int k = 0; k < ncat && ncv; k++) {
      const char *v = cv[ncv == 1 ? 0 : k];
      cnvt[k] = (v[0] == 'T' || v[0] == 't' || v[0] == '1' || v[0] == 'Y' || v[0] == 'y');
      any_cnvt |= cnvt[k];
    }
    free_list(cv, ncv);
    if (any_cnvt) {
      zfile = kv_get("Z_CMVDST_FILE");
      if (!zfile) {
        double v;
        cosmo.omega_m = get_double("OMEGA_M", &v) ? NAN : v;
        cosmo.omega_l = get_double("OMEGA_LAMBDA", &v) ? NAN : v;
        cosmo.w = get_double("DE_EOS_W", &v) ? NAN : v;
        cosmo.err = get_double("CMVDST_ERR", &v) ? NAN : v;
        if (isnan(cosmo.omega_m)) { P_ERR("OMEGA_M is not set\n"); FAIL(FCFC_ERR_CFG); }
        if (fcfc_cosmology_check(&cosmo)) {
          P_ERR("invalid cosmology for COORD_CONVERT (OMEGA_M in (0,1], OMEGA_LAMBDA >= 0, "
                "DE_EOS_W <= -1/3, positive H^2, CMVDST_ERR >= machine epsilon)\n");
          FAIL(FCFC_ERR_CFG);
        }
      }
    }
  }
#else
  (void) zfile; (void) any_cnvt;
  if (get_bool("COORD_CONVERT", 0)) { P_ERR("COORD_CONVERT applies to survey catalogues (FCFC_2PT)\n"); FAIL(FCFC_ERR_CFG); }
#endif
  if (npc < 1 || npco != npc) { P_ERR("PAIR_COUNT and PAIR_COUNT_FILE must be set with equal lengths\n"); FAIL(FCFC_ERR_CFG); }
  if (ncf != ncfo) { P_ERR("CF_ESTIMATOR and CF_OUTPUT_FILE must have equal lengths\n"); FAIL(FCFC_ERR_CFG); }
  const int bintype = get_int("BINNING_SCHEME", 0);
  const int ofmt = get_int("OUTPUT_FORMAT", 0);
  const int overwrite = get_int("OVERWRITE", 0);
  /* DATA_STRUCT: 0 (k-d tree) or 1 (ball tree) in the reference
   * (load_conf.c:825-834); both give the same counts, so the choice has no
   * effect here */
  {
    const int ds = get_int("DATA_STRUCT", 0);
    if (ds != 0 && ds != 1) { P_ERR("invalid DATA_STRUCT: %d\n", ds); FAIL(FCFC_ERR_CFG); }
  }
  const int prec = get_int("GPU_PRECISION", 0);
  const int want_wp = get_bool("PROJECTED_CF", 0);

  /* box */
  double bsize[3] = {0, 0, 0};
#if FCFC_CLI_GEOM == 0
  {
    char **bl; int nb = split_list(kv_get("BOX_SIZE"), &bl);
    if (nb != 1 && nb != 3) { P_ERR("BOX_SIZE must have 1 or 3 values\n"); free_list(bl, nb); FAIL(FCFC_ERR_CFG); }
    for (int k = 0; k < 3; k++) bsize[k] = atof(bl[nb == 1 ? 0 : k]);
    free_list(bl, nb);
  }
#endif
  /* bins */
  {
    int nmu = get_int("MU_BIN_NUM", 1);
    const char *sf = kv_get("SEP_BIN_FILE");
    const char *pf = kv_get("PI_BIN_FILE");
    int e;
    if (sf || pf) {
      double *se = NULL, *pe = NULL; int nse = 0, npe = 0;
      const char *files[2] = {sf, pf};
      for (int f = 0; f < 2; f++) {
        if (!files[f]) continue;
        FILE *fp = fopen(files[f], "r");
        if (!fp) { P_ERR("cannot open `%s'\n", files[f]); FAIL(FCFC_ERR_FILE); }
        char ln[1024]; double *edges = NULL; int ne = 0;
        while (fgets(ln, sizeof ln, fp)) {
          char *p = trim(ln);
          if (!*p || *p == '#') continue;
          double lo, hi;
          if (sscanf(p, "%lf %lf", &lo, &hi) != 2) continue;
          edges = realloc(edges, sizeof(double) * (ne + 2));
          if (ne == 0) edges[ne++] = lo;
          edges[ne++] = hi;
        }
        fclose(fp);
        if (f == 0) { se = edges; nse = ne; } else { pe = edges; npe = ne; }
      }
      if (!sf) {
        double smin, smax, ds;
        if (get_double("SEP_BIN_MIN", &smin) || get_double("SEP_BIN_MAX", &smax) || get_double("SEP_BIN_SIZE", &ds)) { P_ERR("separation bins not set\n"); FAIL(FCFC_ERR_CFG); }
        nse = (int) floor((smax - smin) / ds + 0.5) + 1;
        se = malloc(sizeof(double) * nse);
        for (int i = 0; i < nse; i++) se[i] = smin + ds * i;
      }
      if (bintype == FCFC_BIN_SPI && !pf) {
        double pmin, pmax, dpi;
        if (get_double("PI_BIN_MIN", &pmin) || get_double("PI_BIN_MAX", &pmax) || get_double("PI_BIN_SIZE", &dpi)) { P_ERR("pi bins not set\n"); FAIL(FCFC_ERR_CFG); }
        npe = (int) floor((pmax - pmin) / dpi + 0.5) + 1;
        pe = malloc(sizeof(double) * npe);
        for (int i = 0; i < npe; i++) pe[i] = pmin + dpi * i;
      }
      e = fcfc_bins_edges(&bins, FCFC_CLI_GEOM, bintype, se, nse - 1, nmu, pe, npe - 1,
          FCFC_CLI_GEOM == 0 ? bsize : NULL);
      free(se); free(pe);
    }
    else {
      double smin, smax, ds, pmin = 0, pmax = 0, dpi = 1;
      if (get_double("SEP_BIN_MIN", &smin) || get_double("SEP_BIN_MAX", &smax) || get_double("SEP_BIN_SIZE", &ds)) { P_ERR("separation bins not set\n"); FAIL(FCFC_ERR_CFG); }
      if (bintype == FCFC_BIN_SPI && (get_double("PI_BIN_MIN", &pmin) || get_double("PI_BIN_MAX", &pmax) || get_double("PI_BIN_SIZE", &dpi))) { P_ERR("pi bins not set\n"); FAIL(FCFC_ERR_CFG); }
      e = fcfc_bins_linear(&bins, FCFC_CLI_GEOM, bintype, smin, smax, ds, nmu, pmin, pmax, dpi,
          FCFC_CLI_GEOM == 0 ? bsize : NULL);
    }
    if (e) { P_ERR("invalid binning configuration (%d)\n", e); FAIL(FCFC_ERR_CFG); }
  }
  double bres[3] = {bsize[0] * bins.rescale, bsize[1] * bins.rescale, bsize[2] * bins.rescale};

  /* labels */
  char lab[26];
  for (int i = 0; i < ncat; i++) lab[i] = (i < nlab && labels[i][0]) ? labels[i][0] : (char) ('A' + i);
  cat = calloc(ncat, sizeof(cat_t));
  cnt = calloc(npc, sizeof(double *));
  ncnt = calloc(npc, sizeof(double *));
  norm = calloc(npc, sizeof(double));
  weighted = calloc(npc, sizeof(int));
  pclab = calloc(npc, 3);
  pci[0] = calloc(npc, sizeof(int)); pci[1] = calloc(npc, sizeof(int));
  for (int i = 0; i < npc; i++) {
    if (strlen(pcs[i]) != 2) { P_ERR("invalid PAIR_COUNT `%s'\n", pcs[i]); FAIL(FCFC_ERR_CFG); }
    pclab[i][0] = pcs[i][0]; pclab[i][1] = pcs[i][1];
    for (int s = 0; s < 2; s++) {
      pci[s][i] = -1;
      for (int k = 0; k < ncat; k++) if (lab[k] == pcs[i][s]) pci[s][i] = k;
    }
  }
  /* existing outputs (load_conf.c check_output): OVERWRITE <= 0 aborts,
   * 1 reads existing pair-count files and rewrites the correlation function
   * files, >= 2 rewrites everything.  (The reference asks on the terminal
   * for negative values; here they abort like 0.) */
  comp_pc = calloc(npc, sizeof(int));
  for (int i = 0; i < npc; i++) {
    comp_pc[i] = 1;
    if (!file_exists(pcout[i])) continue;
    if (overwrite <= 0) { P_ERR("abort to avoid overwriting PAIR_COUNT_FILE: `%s'\n", pcout[i]); FAIL(FCFC_ERR_FILE); }
    if (overwrite == 1) comp_pc[i] = 0;
  }
  {
    char **outs[3] = {cfout, mpout, wpout};
    const int nout[3] = {ncfo, (bintype == FCFC_BIN_SMU && npole) ? nmpo : 0,
                         (bintype == FCFC_BIN_SPI && want_wp) ? nwpo : 0};
    const char *keys[3] = {"CF_OUTPUT_FILE", "MULTIPOLE_FILE", "PROJECTED_FILE"};
    for (int q = 0; q < 3; q++)
      for (int i = 0; i < nout[q]; i++)
        if (file_exists(outs[q][i]) && overwrite <= 0) {
          P_ERR("abort to avoid overwriting %s: `%s'\n", keys[q], outs[q][i]);
          FAIL(FCFC_ERR_FILE);
        }
  }
  /* MULTIPOLE: sorted, duplicates removed, 0 <= ell <= 6 (load_conf.c) */
  if (npole > 1) {
    for (int i = 1; i < npole; i++)
      for (int j = i; j > 0 && atoi(poles[j - 1]) > atoi(poles[j]); j--) {
        char *t = poles[j]; poles[j] = poles[j - 1]; poles[j - 1] = t;
      }
    int m = 0;
    for (int i = 1; i < npole; i++)
      if (atoi(poles[i]) != atoi(poles[m])) { m++; char *t = poles[m]; poles[m] = poles[i]; poles[i] = t; }
    for (int i = m + 1; i < npole; i++) { free(poles[i]); poles[i] = NULL; }
    npole = m + 1;
  }
  for (int l = 0; l < npole; l++)
    if (atoi(poles[l]) < 0 || atoi(poles[l]) > 6) { P_ERR("MULTIPOLE must be between 0 and 6\n"); FAIL(FCFC_ERR_CFG); }
  /* catalogue labels are needed for the pair counts that are computed */
  for (int i = 0; i < npc; i++)
    if (comp_pc[i] && (pci[0][i] < 0 || pci[1][i] < 0)) {
      P_ERR("catalog label not found for PAIR_COUNT: %s\n", pcs[i]); FAIL(FCFC_ERR_CFG);
    }
  if (bintype == FCFC_BIN_SMU && npole && ncf && nmpo != ncf) {
    P_ERR("MULTIPOLE_FILE must have as many entries as CF_ESTIMATOR\n"); FAIL(FCFC_ERR_CFG);
  }
  has_wt = calloc(ncat, sizeof(int));
  for (int k = 0; k < ncat; k++) has_wt[k] = (k < nwt && wts[k][0] && strcmp(wts[k], "1"));
  if (fcfc_gpu_init(&ctx, 0)) { P_ERR("no usable GPU\n"); FAIL(FCFC_ERR_UNKNOWN); }

  for (int i = 0; i < npc; i++) {
    const int a = pci[0][i], b = pci[1][i], isauto = (a == b);
    cnt[i] = calloc(bins.ntot, sizeof(double));
    ncnt[i] = calloc(bins.ntot, sizeof(double));
    /* an existing pair-count file with OVERWRITE = 1: its normalised counts
     * are read instead of counted (eval_cf.c:64-85) */
    if (!comp_pc[i]) {
      printf("Reading %s pairs ...\n", pcs[i]);
      if (verbose) printf("  Filename: %s\n", pcout[i]);
      if ((ret = read_pair_file(pcout[i], &bins, bres, ncnt[i]))) goto done;
      printf("[DONE]\n");
      continue;
    }
    weighted[i] = has_wt[a] || has_wt[b];
    for (int s = 0; s < 2; s++) {
      int k = pci[s][i];
      if (cat[k].loaded) continue;
      /* POSITION, WEIGHT and SELECTION expressions of catalogue k */
      xprog_t xp[4], xs;
      int nc = has_wt[k] ? 4 : 3, hs = (k < nsel && sels[k][0]);
      memset(xp, 0, sizeof xp);
      memset(&xs, 0, sizeof xs);
      int ce = 0;
      const int fits = (ntype == 1) ? atoi(types[0]) == 1 : (ntype > k && atoi(types[k]) == 1);
      fits_t ft;
      memset(&ft, 0, sizeof ft);
      if (fits) ce = fits_open(cats[k], &ft);
      for (int d = 0; d < 3 && !ce; d++) ce = x_compile_named(pos[3 * k + d], &xp[d], ft.names, ft.ncol);
      if (!ce && nc == 4) ce = x_compile_named(wts[k], &xp[3], ft.names, ft.ncol);
      if (!ce && hs) ce = x_compile_named(sels[k], &xs, ft.names, ft.ncol);
      printf("Construct the grid for catalog '%c' ...\n", lab[k]);
      long skip = (k < nskip) ? atol(skips[k]) : 0;
      char cm = (k < ncmt && cmts[k][0]) ? cmts[k][0] : '\0';
      if (!ce && fits) ce = read_fits(cats[k], &ft, xp, nc, hs ? &xs : NULL, &cat[k]);
      else if (!ce) ce = read_ascii(cats[k], skip, cm, fmts[k], xp, nc, hs ? &xs : NULL, &cat[k]);
      fits_close(&ft);
      for (int d = 0; d < 4; d++) free(xp[d].code);
      free(xs.code);
      if (ce) { ret = (ce == FCFC_ERR_AST) ? FCFC_ERR_CFG : ce; goto done; }
      if (verbose) printf("  Number of recorded objects: %lld\n", (long long) cat[k].n);
#if FCFC_CLI_GEOM == 1
      /* (RA, Dec, z) -> comoving Cartesian (cnvt_coord.c:cnvt_coord) */
      if (cnvt && cnvt[k]) {
        int e2, order = 0;
        if (zfile) {
          double *zs = NULL, *ds = NULL;
          int64_t ns = 0;
          if ((e2 = read_table2(zfile, &zs, &ds, &ns))) { free(zs); free(ds); ret = e2; goto done; }
          e2 = fcfc_convert_coords_table(zs, ds, ns, cat[k].n, cat[k].x, cat[k].y, cat[k].z);
          free(zs); free(ds);
        }
        else e2 = fcfc_convert_coords(&cosmo, cat[k].n, cat[k].x, cat[k].y, cat[k].z, &order);
        if (e2) { P_ERR("coordinate conversion of catalog '%c' failed (%d)\n", lab[k], e2); ret = FCFC_ERR_CNVT; goto done; }
        if (verbose) {
          if (zfile) printf("  Coordinates converted using cubic spline interpolation\n");
          else printf("  Coordinates converted using Legendre-Gauss integration with order %d\n", order);
        }
      }
#endif
    }
    /* catalogue preparation (rescale, survey |r|^2) */
    fcfc_cat ca, cb;
    double *sv[2] = {NULL, NULL};
    for (int s = 0; s < 2; s++) {
      int k = pci[s][i];
      cat_t *c = &cat[k];
      if (c->loaded == 1) {
        if (fcfc_prepare_catalog(&bins, c->n, c->x, c->y, c->z, NULL, c->w, &c->sumw)) FAIL(FCFC_ERR_DATA);
        c->loaded = 2;   /* rescaled once */
      }
      if (FCFC_CLI_GEOM == 1 && bintype != FCFC_BIN_ISO) {
        sv[s] = malloc(sizeof(double) * (c->n ? c->n : 1));
        fcfc_bins nb = bins; nb.rescale = 1;
        fcfc_prepare_catalog(&nb, c->n, c->x, c->y, c->z, sv[s], NULL, NULL);
      }
      fcfc_cat *t = s ? &cb : &ca;
      t->n = c->n; t->x = c->x; t->y = c->y; t->z = c->z; t->s = sv[s];
      t->w = weighted[i] ? c->w : NULL; t->on_device = 0;
    }
    printf("Counting %c%c pairs ...\n", pclab[i][0], pclab[i][1]);
    fflush(stdout);
    uint32_t flags = (isauto ? FCFC_GPU_AUTO : 0) | (weighted[i] ? FCFC_GPU_WEIGHTED : 0) |
        (prec ? FCFC_GPU_PREC_FP32 : 0);
    int e = fcfc_gpu_count(ctx, &ca, &cb, &bins, FCFC_CLI_GEOM == 0 ? bres : NULL, flags, cnt[i]);
    free(sv[0]); free(sv[1]);
    if (e) { P_ERR("GPU pair counting failed: %s\n", fcfc_gpu_last_error(ctx)); FAIL(e); }
    /* caller conventions of eval_pairs (eval_cf.c:117-167) */
    const cat_t *A = &cat[a], *B = &cat[b];
    double wa = has_wt[a] ? A->sumw : (double) A->n, wb = has_wt[b] ? B->sumw : (double) B->n;
    if (isauto) {
      if (weighted[i]) { for (int k = 0; k < bins.ntot; k++) cnt[i][k] *= 2; norm[i] = wa * (wa - 1); }
      else { for (int k = 0; k < bins.ntot; k++) ((int64_t *) cnt[i])[k] *= 2; norm[i] = (double) A->n * (A->n - 1); }
    }
    else norm[i] = weighted[i] ? wa * wb : (double) A->n * B->n;
    for (int k = 0; k < bins.ntot; k++)
      ncnt[i][k] = (weighted[i] ? cnt[i][k] : (double) ((int64_t *) cnt[i])[k]) / norm[i];
    if ((e = fcfc_save_pair_count(pcout[i], &bins, ofmt, bres, isauto, weighted[i], A->n, B->n, wa, wb, cnt[i]))) {
      P_ERR("cannot write `%s'\n", pcout[i]); FAIL(e);
    }
    if (verbose) printf("  Results written to file: `%s'\n", pcout[i]);
    printf("[DONE]\n");
  }

  /* correlation functions (eval_cf.c:258-388) */
  if (ncf) {
    printf("Evaluate correlation function estimators ...\n");
    double *rr = NULL;
#if FCFC_CLI_GEOM == 0
    rr = calloc(bins.ntot, sizeof(double));
    if (bins.bintype == FCFC_BIN_SPI) {
      double fac = 2 * M_PI / (bsize[0] * bsize[1] * bsize[2]);
      for (int j = 0; j < bins.np; j++) {
        double dpi = bins.pbin_raw[j + 1] - bins.pbin_raw[j];
        for (int k = 0; k < bins.ns; k++)
          rr[k + j * bins.ns] = fac * dpi * (bins.sbin_raw[k + 1] * bins.sbin_raw[k + 1] - bins.sbin_raw[k] * bins.sbin_raw[k]);
      }
    }
    else {
      double fac = 4 * M_PI / (3 * bins.nmu * bsize[0] * bsize[1] * bsize[2]);
      for (int k = 0; k < bins.ns; k++) {
        double v = fac * (pow(bins.sbin_raw[k + 1], 3) - pow(bins.sbin_raw[k], 3));
        for (int j = 0; j < bins.nmu; j++) rr[k + j * bins.ns] = v;
      }
    }
#endif
    double *vals = malloc(sizeof(double) * npc);
    double *xi = malloc(sizeof(double) * bins.ntot);
    for (int c = 0; c < ncf; c++) {
      for (int k = 0; k < bins.ntot; k++) {
        for (int p = 0; p < npc; p++) vals[p] = ncnt[p][k];
        ex_t ex = {cfs[c], vals, pclab, npc, rr ? rr[k] : NAN, 0};
        xi[k] = ex_sum(&ex);
        if (ex.err || *ex.p) { P_ERR("cannot evaluate CF_ESTIMATOR `%s'\n", cfs[c]); free(vals); free(xi); free(rr); FAIL(FCFC_ERR_CFG); }
      }
      FILE *fp = fopen(cfout[c], "w");
      if (!fp) { P_ERR("cannot write `%s'\n", cfout[c]); free(vals); free(xi); free(rr); FAIL(FCFC_ERR_FILE); }
      if (bins.bintype == FCFC_BIN_SMU) {
        fprintf(fp, "# Columns: s_min(1) s_max(2) mu_min(3) mu_max(4) xi(5)\n");
        for (int j = 0; j < bins.nmu; j++)
          for (int k = 0; k < bins.ns; k++)
            fprintf(fp, "%.10lg %.10lg %.10lg %.10lg %.10lg\n", bins.sbin_raw[k], bins.sbin_raw[k + 1],
                j / (double) bins.nmu, (j + 1) / (double) bins.nmu, xi[k + j * bins.ns]);
      }
      else if (bins.bintype == FCFC_BIN_SPI) {
        fprintf(fp, "# Columns: s_perp_min(1) s_perp_max(2) pi_min(3) pi_max(4) xi(5)\n");
        for (int j = 0; j < bins.np; j++)
          for (int k = 0; k < bins.ns; k++)
            fprintf(fp, "%.10lg %.10lg %.10lg %.10lg %.10lg\n", bins.sbin_raw[k], bins.sbin_raw[k + 1],
                bins.pbin_raw[j], bins.pbin_raw[j + 1], xi[k + j * bins.ns]);
      }
      else {
        fprintf(fp, "# Columns: s_cen(1) s_min(2) s_max(3) xi_iso(4)\n");
        for (int k = 0; k < bins.ns; k++)
          fprintf(fp, "%.10lg %.10lg %.10lg %.10lg\n", (bins.sbin_raw[k] + bins.sbin_raw[k + 1]) * 0.5,
              bins.sbin_raw[k], bins.sbin_raw[k + 1], xi[k]);
      }
      fclose(fp);
      /* multipoles (eval_cf.c:326-354) */
      if (bins.bintype == FCFC_BIN_SMU && npole && c < nmpo) {
        fp = fopen(mpout[c], "w");
        if (!fp) { free(vals); free(xi); free(rr); FAIL(FCFC_ERR_FILE); }
        fprintf(fp, "# Columns: s_cen(1) s_min(2) s_max(3)");
        for (int l = 0; l < npole; l++) fprintf(fp, " xi_%d(%d)", atoi(poles[l]), 4 + l);
        fprintf(fp, "\n");
        for (int k = 0; k < bins.ns; k++) {
          fprintf(fp, "%.10lg %.10lg %.10lg", (bins.sbin_raw[k] + bins.sbin_raw[k + 1]) * 0.5,
              bins.sbin_raw[k], bins.sbin_raw[k + 1]);
          for (int l = 0; l < npole; l++) {
            int ell = atoi(poles[l]);
            double fac = (2 * ell + 1) / (double) bins.nmu, v = 0;
            for (int j = 0; j < bins.nmu; j++) v += xi[k + j * bins.ns] * fac * legp(ell, (j + 0.5) / (double) bins.nmu);
            fprintf(fp, " %.10lg", v);
          }
          fprintf(fp, "\n");
        }
        fclose(fp);
      }
      /* projected CF (eval_cf.c:365-388) */
      if (bins.bintype == FCFC_BIN_SPI && want_wp && c < nwpo) {
        fp = fopen(wpout[c], "w");
        if (!fp) { free(vals); free(xi); free(rr); FAIL(FCFC_ERR_FILE); }
        fprintf(fp, "# Columns: s_perp_cen(1) s_perp_min(2) s_perp_max(3) wp(4)\n");
        for (int k = 0; k < bins.ns; k++) {
          double v = 0;
          for (int j = 0; j < bins.np; j++) v += 2 * xi[k + j * bins.ns] * (bins.pbin_raw[j + 1] - bins.pbin_raw[j]);
          fprintf(fp, "%.10lg %.10lg %.10lg %.10lg\n", (bins.sbin_raw[k] + bins.sbin_raw[k + 1]) * 0.5,
              bins.sbin_raw[k], bins.sbin_raw[k + 1], v);
        }
        fclose(fp);
      }
    }
    free(vals); free(xi); free(rr);
    printf("[DONE]\n");
  }

done:
  if (ctx) fcfc_gpu_destroy(ctx);
  fcfc_bins_free(&bins);
  if (cat) for (int k = 0; k < ncat; k++) { free(cat[k].x); free(cat[k].y); free(cat[k].z); free(cat[k].w); }
  free(cat);
  if (cnt) for (int i = 0; i < npc; i++) { free(cnt[i]); free(ncnt[i]); }
  free(cnt); free(ncnt); free(norm); free(weighted); free(comp_pc); free(cnvt);
  free(has_wt); free(pclab); free(pci[0]); free(pci[1]);
  free_list(sels, nsel);
  free_list(cats, ncat); free_list(labels, nlab); free_list(pos, npos); free_list(wts, nwt);
  free_list(fmts, nfmt); free_list(pcs, npc); free_list(pcout, npco); free_list(cfs, ncf);
  free_list(cfout, ncfo); free_list(mpout, nmpo); free_list(wpout, nwpo); free_list(poles, npole);
  free_list(skips, nskip); free_list(cmts, ncmt); free_list(types, ntype);
  return ret;
}
