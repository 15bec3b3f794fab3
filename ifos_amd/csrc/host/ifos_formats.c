/* ifos_formats.c -- the reference's parameter, model and seismogram file
 * formats (include/ifos2d_formats.h). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ifos2d_formats.h"

void ifos_set_error(const char *fmt, ...);

#define LINE_MAX_CHARS 4096

struct ifos_json {
    int n, cap;
    char **name, **value;
};

/* the first blank-delimited word of s (sscanf "%s" in
 * remove_blankspaces_around_string, src/json_parser.c:402-415) */
static char *first_word(const char *s, size_t len)
{
    size_t a = 0;
    while (a < len && (s[a] == ' ' || s[a] == '\t' || s[a] == '\n' || s[a] == '\r')) a++;
    size_t b = a;
    while (b < len && !(s[b] == ' ' || s[b] == '\t' || s[b] == '\n' || s[b] == '\r')) b++;
    char *w = (char *)malloc(b - a + 1);
    memcpy(w, s + a, b - a);
    w[b - a] = '\0';
    return w;
}

static int json_add(struct ifos_json *j, char *name, char *value)
{
    if (j->n == j->cap) {
        j->cap = j->cap ? 2 * j->cap : 64;
        j->name = (char **)realloc(j->name, sizeof(char *) * j->cap);
        j->value = (char **)realloc(j->value, sizeof(char *) * j->cap);
    }
    j->name[j->n] = name;
    j->value[j->n] = value;
    j->n++;
    return 0;
}

static int count_char(const char *s, char c)
{
    int n = 0;
    for (; *s; s++) n += *s == c;
    return n;
}

/* split the text between two quotes at commas into up to five words */
static int split_commas(const char *s, size_t len, char **out, int max)
{
    int n = 0;
    size_t a = 0;
    for (size_t i = 0; i <= len; i++)
        if (i == len || s[i] == ',') {
            if (n == max) return -1;
            out[n++] = first_word(s + a, i - a);
            a = i + 1;
        }
    return n;
}

int ifos_json_load(const char *path, ifos_json **out)
{
    *out = NULL;
    FILE *fp = fopen(path, "r");
    if (!fp) return ifos_set_error("could not open input file '%s'", path), -2;
    struct ifos_json *j = (struct ifos_json *)calloc(1, sizeof *j);
    char line[LINE_MAX_CHARS];
    int lineno = 0;
    while (fgets(line, sizeof line, fp)) {
        lineno++;
        /* an object line has a colon and double quotes and is no comment
         * (src/json_parser.c:59-60) */
        if (!(strchr(line, ':') && strchr(line, '"')) || strstr(line, "comment") || strstr(line, "Comment")) continue;
        if (count_char(line, '"') != 4) {
            fclose(fp);
            ifos_json_free(j);
            return ifos_set_error("%s:%d: an object line needs two quoted strings", path, lineno), -2;
        }
        const char *q1 = strchr(line, '"'), *q2 = strchr(q1 + 1, '"');
        const char *q3 = strchr(q2 + 1, '"'), *q4 = strchr(q3 + 1, '"');
        const char *colon = strchr(q2 + 1, ':');
        if (!colon || colon > q3) {
            fclose(fp);
            ifos_json_free(j);
            return ifos_set_error("%s:%d: cannot read object name and object value", path, lineno), -2;
        }
        char *names[5] = {0}, *values[5] = {0};
        const int nn = split_commas(q1 + 1, (size_t)(q2 - q1 - 1), names, 5);
        const int nv = split_commas(q3 + 1, (size_t)(q4 - q3 - 1), values, 5);
        if (nn < 1 || nn != nv) {
            for (int k = 0; k < 5; k++) {
                free(names[k]);
                free(values[k]);
            }
            fclose(fp);
            ifos_json_free(j);
            return ifos_set_error("%s:%d: %d object names but %d values", path, lineno, nn, nv), -2;
        }
        for (int k = 0; k < nn; k++) json_add(j, names[k], values[k]);
    }
    fclose(fp);
    *out = j;
    return 0;
}

void ifos_json_free(ifos_json *j)
{
    if (!j) return;
    for (int k = 0; k < j->n; k++) {
        free(j->name[k]);
        free(j->value[k]);
    }
    free(j->name);
    free(j->value);
    free(j);
}

int ifos_json_count(const ifos_json *j) { return j->n; }
const char *ifos_json_name(const ifos_json *j, int k) { return k >= 0 && k < j->n ? j->name[k] : NULL; }
const char *ifos_json_value(const ifos_json *j, int k) { return k >= 0 && k < j->n ? j->value[k] : NULL; }

const char *ifos_json_get(const ifos_json *j, const char *name)
{
    for (int k = 0; k < j->n; k++)
        if (strcmp(j->name[k], name) == 0) return j->value[k];
    return NULL;
}

int ifos_json_float(const ifos_json *j, const char *name, float *out)
{
    const char *v = ifos_json_get(j, name);
    if (!v) return 1;
    char *end = NULL;
    const double d = strtod(v, &end);
    if (!*v || end == v || *end) return ifos_set_error("value of %s is not a number: '%s'", name, v), -2;
    *out = (float)d;
    return 0;
}

int ifos_json_int(const ifos_json *j, const char *name, int *out)
{
    const char *v = ifos_json_get(j, name);
    if (!v) return 1;
    char *end = NULL;
    const double d = strtod(v, &end);
    double ip;
    if (!*v || end == v || *end || modf(d, &ip) != 0.0)
        return ifos_set_error("value of %s is not an integer: '%s'", name, v), -2;
    *out = (int)d;
    return 0;
}

int ifos_config_from_json(const char *path, ifos_config *c)
{
    ifos_json *j = NULL;
    int rc = ifos_json_load(path, &j);
    if (rc) return rc;
    struct { const char *k; int *p; int required; } ints[] = {
        {"NX", &c->NX, 1}, {"NY", &c->NY, 1}, {"FDORDER", &c->FDORDER, 1}, {"MAXRELERROR", &c->MAXRELERROR, 0},
        {"FW", &c->FW, 0}, {"FREE_SURF", &c->FREE_SURF, 0}, {"BOUNDARY", &c->BOUNDARY, 0}, {"L", &c->L, 0},
        {"VELOCITY", &c->VELOCITY, 0}, {"ADJOINT_TYPE", &c->ADJOINT_TYPE, 0},
        {"PARAMETERIZATION", &c->PARAMETERIZATION, 0}, {"DTINV", &c->DTINV, 0},
        {"SOURCE_SHAPE", &c->SOURCE_SHAPE, 0}, {"SOURCE_TYPE", &c->SOURCE_TYPE, 0}, {"TRKILL", &c->TRKILL, 0}};
    struct { const char *k; float *p; int required; } flts[] = {
        {"DH", &c->DH, 1}, {"DT", &c->DT, 1}, {"VPPML", &c->VPPML, 0}, {"FPML", &c->FPML, 0},
        {"npower", &c->npower, 0}, {"k_max_PML", &c->k_max_PML, 0}, {"TAU", &c->TAU, 0}, {"F_REF", &c->F_REF, 0}};
    for (size_t k = 0; k < sizeof ints / sizeof ints[0]; k++) {
        const int r = ifos_json_int(j, ints[k].k, ints[k].p);
        if (r < 0 || (r == 1 && ints[k].required)) {
            if (r == 1) ifos_set_error("variable %s could not be retrieved from %s", ints[k].k, path);
            ifos_json_free(j);
            return -2;
        }
    }
    for (size_t k = 0; k < sizeof flts / sizeof flts[0]; k++) {
        const int r = ifos_json_float(j, flts[k].k, flts[k].p);
        if (r < 0 || (r == 1 && flts[k].required)) {
            if (r == 1) ifos_set_error("variable %s could not be retrieved from %s", flts[k].k, path);
            ifos_json_free(j);
            return -2;
        }
    }
    float TIME = 0.0f;
    rc = ifos_json_float(j, "TIME", &TIME);
    if (rc) {
        if (rc == 1) ifos_set_error("variable TIME could not be retrieved from %s", path);
        ifos_json_free(j);
        return -2;
    }
    c->NT = (int)floor(TIME / c->DT + 0.5);  /* iround, src/fd.h:15 */
    /* FL1..FL_L of the relaxation mechanisms (src/read_par_json.c:387-412) */
    static const char *fl[5] = {"FL1", "FL2", "FL3", "FL4", "FL5"};
    for (int l = 0; l < c->L && l < 5; l++) {
        rc = ifos_json_float(j, fl[l], &c->FL[l + 1]);
        if (rc) {
            if (rc == 1) ifos_set_error("variable %s could not be retrieved from %s", fl[l], path);
            ifos_json_free(j);
            return -2;
        }
    }
    ifos_json_free(j);
    return 0;
}

int ifos_model_read(const char *path, int NX, int NY, float *m)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return ifos_set_error("could not open model file '%s'", path), -2;
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
            if (fread(&m[(size_t)j * NX + i], sizeof(float), 1, fp) != 1) {
                fclose(fp);
                return ifos_set_error("model file '%s' ends before %d x %d values", path, NX, NY), -2;
            }
    fclose(fp);
    return 0;
}

int ifos_model_write(const char *path, int NX, int NY, const float *m)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return ifos_set_error("could not open '%s' for writing", path), -2;
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) fwrite(&m[(size_t)j * NX + i], sizeof(float), 1, fp);
    const int bad = ferror(fp);
    fclose(fp);
    return bad ? (ifos_set_error("write to '%s' failed", path), -1) : 0;
}

/* SEG-Y trace header byte offsets (the SU segy.h layout) */
enum {
    SU_TRACL = 0, SU_CDP = 20, SU_TRID = 28, SU_OFFSET = 36, SU_GELEV = 40, SU_SDEPTH = 48, SU_SWDEP = 60,
    SU_SCALEL = 68, SU_SCALCO = 70, SU_SX = 72, SU_GX = 80, SU_NS = 114, SU_DT = 116, SU_D1 = 180, SU_HDR = 240
};

static void put_i32(unsigned char *h, int off, int32_t v) { memcpy(h + off, &v, 4); }
static void put_i16(unsigned char *h, int off, int16_t v) { memcpy(h + off, &v, 2); }
static int iround_d(double x) { return (int)floor(x + 0.5); }

int ifos_su_write(const char *path, int ntr, int ns, const float *data, const int *recpos, float DH, float DT,
                  int ndt, float xs, float ys, float refrec_x, float refrec_y)
{
    if (ns < 1 || ns > 65535) return ifos_set_error("SU traces hold 1..65535 samples (got %d)", ns), -2;
    FILE *fp = fopen(path, "wb");
    if (!fp) return ifos_set_error("could not open '%s' for writing", path), -2;
    const float xshift = 800.0f, yshift = 800.0f;  /* src/outseis_glob.c:33 */
    const double PI = 3.1415926535897932384626433832795;
    for (int t = 0; t < ntr; t++) {
        unsigned char h[SU_HDR];
        memset(h, 0, sizeof h);
        const float xr = recpos[2 * t] * DH, yr = recpos[2 * t + 1] * DH;
        const float x = xr - refrec_x, y = yr - refrec_y;
        put_i32(h, SU_TRACL, t + 1);
        put_i32(h, SU_CDP, t + 1);
        put_i16(h, SU_TRID, 1);
        put_i32(h, SU_OFFSET, iround_d(sqrt((xs - xr) * (xs - xr) + (ys - yr) * (ys - yr)) * 1000.0));
        put_i32(h, SU_GELEV, iround_d(yr * 1000.0));
        put_i32(h, SU_SDEPTH, iround_d(ys * 1000.0));
        put_i32(h, SU_SWDEP, iround_d(((360.0 / (2.0 * PI)) * atan2(x - xshift, y - yshift)) * 1000.0));
        put_i16(h, SU_SCALEL, -1000);
        put_i16(h, SU_SCALCO, -1000);
        put_i32(h, SU_SX, iround_d(xs * 1000.0));
        put_i32(h, SU_GX, iround_d(xr * 1000.0));
        const uint16_t nsu = (uint16_t)ns, dtu = (uint16_t)iround_d(((float)ndt * DT) * 1.0e6);
        memcpy(h + SU_NS, &nsu, 2);
        memcpy(h + SU_DT, &dtu, 2);
        const float d1 = (float)dtu * 1.0e-6;
        memcpy(h + SU_D1, &d1, 4);
        fwrite(h, 1, sizeof h, fp);
        fwrite(data + (size_t)t * ns, sizeof(float), (size_t)ns, fp);
    }
    const int bad = ferror(fp);
    fclose(fp);
    return bad ? (ifos_set_error("write to '%s' failed", path), -1) : 0;
}

int ifos_su_read(const char *path, int max_tr, int ns, float *data, int *ntr_out)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return ifos_set_error("could not open SU file '%s'", path), -2;
    int t = 0;
    unsigned char h[SU_HDR];
    while (fread(h, 1, sizeof h, fp) == sizeof h) {
        uint16_t nsu;
        memcpy(&nsu, h + SU_NS, 2);
        if ((int)nsu != ns) {
            fclose(fp);
            return ifos_set_error("'%s' trace %d has %d samples, expected %d", path, t + 1, (int)nsu, ns), -2;
        }
        if (t == max_tr) {
            fclose(fp);
            return ifos_set_error("'%s' holds more than %d traces", path, max_tr), -2;
        }
        if (fread(data + (size_t)t * ns, sizeof(float), (size_t)ns, fp) != (size_t)ns) {
            fclose(fp);
            return ifos_set_error("'%s' ends inside trace %d", path, t + 1), -2;
        }
        t++;
    }
    fclose(fp);
    *ntr_out = t;
    return 0;
}

/* ---- trace kill tables (src/calc_res.c:66-126, src/create_trkill_table.c) ---- */

int ifos_trkill_read(const char *path, int ntr_glob, int nsrc_glob, int *table)
{
    if (ntr_glob < 0 || nsrc_glob < 0 || !table) return ifos_set_error("ifos_trkill_read: bad arguments"), -2;
    FILE *f = fopen(path, "r");
    if (!f) return ifos_set_error(" Trace kill file could not be opened!"), -2;
    /* receiver-major, one integer per (receiver, shot); running into the
     * end of the file during a read fails as in the reference */
    for (int r = 0; r < ntr_glob; r++)
        for (int s = 0; s < nsrc_glob; s++) {
            int v = 0;
            const int got = fscanf(f, "%d", &v);
            if (feof(f) || got != 1) {
                fclose(f);
                return ifos_set_error(" Error while reading TraceKill file. Check dimensions!"), -2;
            }
            table[(long)r * nsrc_glob + s] = v;
        }
    fclose(f);
    return 0;
}

int ifos_trkill_offsets(int ntr_glob, const int *recpos, int nsrc_glob, const float *srcpos, float DH, float lower,
                        float upper, int *table)
{
    if (ntr_glob < 0 || nsrc_glob < 0 || !table) return ifos_set_error("ifos_trkill_offsets: bad arguments"), -2;
    for (int s = 0; s < nsrc_glob; s++) {
        const float xs = srcpos[2 * s], ys = srcpos[2 * s + 1];
        for (int r = 0; r < ntr_glob; r++) {
            const float dx = xs - recpos[2 * r] * DH, dy = ys - recpos[2 * r + 1] * DH;
            const float offset = (float)fabs(sqrt((double)(dx * dx + dy * dy)));
            if (offset > lower && offset < upper) table[(long)r * nsrc_glob + s] = 1;
        }
    }
    return 0;
}

/* ---- observed traces as inseis reads them (src/inseis.c:70-83) ---------- */

int ifos_su_read_observed(const char *path, int ntr, int ns, float *data)
{
    int got = 0;
    const int rc = ifos_su_read(path, ntr, ns, data, &got);
    if (rc) return rc;
    if (got != ntr) return ifos_set_error("'%s' holds %d traces, expected %d", path, got, ntr), -2;
    /* section[t][1] = 0, section[t][k] = sample k of the file for k = 2..ns */
    for (int t = 0; t < ntr; t++) data[(size_t)t * ns] = 0.0f;
    return 0;
}

/* ---- source and receiver positions (src/sources.c:27, src/receiver.c:27) ---- */

static int iround_f(float x) { return (int)floor(x + 0.5); }

int ifos_sources_read(const char *path, int SOURCE_TYPE, int max_src, float *srcpos, int *nsrc_out, float *fc_max)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return ifos_set_error(" Source file could not be opened !"), -2;
    int nsrc = 0;
    if (fscanf(fp, "%d", &nsrc) != 1 || nsrc < 0) {
        fclose(fp);
        return ifos_set_error("'%s': no source count on the first line", path), -2;
    }
    if (nsrc > max_src) {
        fclose(fp);
        return ifos_set_error("'%s' lists %d sources, room for %d", path, nsrc, max_src), -2;
    }
    rewind(fp);
    char line[256];
    if (!fgets(line, 255, fp)) line[0] = 0; /* the count line */
    float fc = 0.0f;
    for (int l = 0; l < nsrc; l++) {
        float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (!fgets(line, 255, fp)) line[0] = 0;
        /* columns: x, z, y, tshift, fc, amp, azimuth, type (z is read and
         * dropped); missing trailing columns take the defaults of the
         * reference's fall-through switch */
        int n = sscanf(line, "%f%f%f%f%f%f%f%f", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
        if (n < 0) n = 0;
        if (n < 6) {
            fclose(fp);
            return ifos_set_error("Missing parameter in SOURCE_FILE! (source %d of '%s')", l + 1, path), -2;
        }
        if (n < 7) v[6] = 0.0f;
        if (n < 8) v[7] = (float)SOURCE_TYPE;
        float *r = srcpos + (size_t)l * IFOS_SRC_STRIDE;
        r[0] = v[0];        /* x (m) */
        r[1] = v[2];        /* y (m) */
        r[2] = 0.0f;
        r[3] = v[3];        /* tshift */
        r[4] = v[4];        /* fc */
        r[5] = v[5];        /* amp */
        r[6] = v[6];        /* azimuth */
        r[7] = v[7];        /* type */
        fc = r[4];
    }
    fclose(fp);
    for (int l = 0; l < nsrc; l++)
        if (srcpos[(size_t)l * IFOS_SRC_STRIDE + 4] > fc) fc = srcpos[(size_t)l * IFOS_SRC_STRIDE + 4];
    *nsrc_out = nsrc;
    if (fc_max) *fc_max = fc;
    return 0;
}

int ifos_sources_to_grid(int nsrc, const float *srcpos_m, float DH, int NX, int NY, float *srcpos, int *nsrc_out)
{
    int k = 0;
    for (int l = 0; l < nsrc; l++) {
        const float *s = srcpos_m + (size_t)l * IFOS_SRC_STRIDE;
        const int ix = iround_f(s[0] / DH), iy = iround_f(s[1] / DH);
        /* one rank covering the grid: POS = (0,0), IENDX = NX, IENDY = NY;
         * C division and remainder truncate towards zero */
        if ((ix - 1) / NX != 0 || (iy - 1) / NY != 0) continue;
        float *d = srcpos + (size_t)k * IFOS_SRC_STRIDE;
        d[0] = (float)(((ix - 1) % NX) + 1);
        d[1] = (float)(((iy - 1) % NY) + 1);
        d[2] = 0.0f;
        for (int c = 3; c < 8; c++) d[c] = s[c];
        k++;
    }
    *nsrc_out = k;
    return 0;
}

int ifos_receivers_read(const char *path, float DH, float refrec_x, float refrec_y, int max_tr, int *recpos,
                        int *ntr_out)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return ifos_set_error(" Receiver file could not be opened !"), -2;
    int n = 0, c;
    while ((c = fgetc(fp)) != EOF)
        if (c == '\n') ++n;
    rewind(fp);
    int *pos = (int *)malloc(sizeof(int) * 2 * (size_t)(n > 0 ? n : 1));
    if (!pos) {
        fclose(fp);
        return ifos_set_error("out of host memory"), -1;
    }
    for (int t = 0; t < n; t++) {
        float x = 0.0f, y = 0.0f;
        if (fscanf(fp, "%f%f\n", &x, &y) != 2) {
            free(pos);
            fclose(fp);
            return ifos_set_error("'%s': receiver line %d unreadable", path, t + 1), -2;
        }
        pos[2 * t] = iround_f((x + refrec_x) / DH);
        pos[2 * t + 1] = iround_f((y + refrec_y) / DH);
    }
    fclose(fp);
    /* a receiver on the grid point of an earlier one is dropped; entries are
     * kept while their x index is positive (the reference's duplicate
     * marker is a negative x index) */
    int flag = 0;
    for (int a = 0; a < n - 1; a++)
        for (int b = a + 1; b < n; b++)
            if (pos[2 * a] == pos[2 * b] && pos[2 * a + 1] == pos[2 * b + 1]) pos[2 * b] = -(++flag);
    int k = 0;
    for (int t = 0; t < n; t++)
        if (pos[2 * t] > 0) {
            if (k == max_tr) {
                free(pos);
                return ifos_set_error("'%s' lists more than %d receivers", path, max_tr), -2;
            }
            recpos[2 * k] = pos[2 * t];
            recpos[2 * k + 1] = pos[2 * t + 1];
            k++;
        }
    free(pos);
    *ntr_out = k;
    return 0;
}

int ifos_receivers_line(float DH, float XREC1, float YREC1, float XREC2, float YREC2, int NGEOPH, int max_tr,
                        int *recpos, int *ntr_out)
{
    if (NGEOPH < 1) return ifos_set_error("NGEOPH must be positive"), -2;
    const int x1 = iround_f(XREC1 / DH), y1 = iround_f(YREC1 / DH), x2 = iround_f(XREC2 / DH),
              y2 = iround_f(YREC2 / DH);
    int ntr;
    /* nzrec1 = nzrec2 = 1 in the reference: horizontal unless the line is
     * steeper than 45 degrees (src/receiver.c:172-204); integer slopes */
    if (abs(y2 - y1) <= abs(x2 - x1)) {
        if (x2 == x1) return ifos_set_error("receiver line of one point"), -2;
        ntr = (x2 - x1) / NGEOPH + 1;
        if (ntr > max_tr) return ifos_set_error("%d receivers, room for %d", ntr, max_tr), -2;
        for (int x = x1; x <= x2; x += NGEOPH) {
            const int itr = (x - x1) / NGEOPH;
            recpos[2 * itr] = x;
            recpos[2 * itr + 1] = y1 + ((y2 - y1) / (x2 - x1) * (x - x1));
        }
    } else {
        ntr = (y2 - y1) / NGEOPH + 1;
        if (ntr > max_tr) return ifos_set_error("%d receivers, room for %d", ntr, max_tr), -2;
        for (int y = y1; y <= y2; y += NGEOPH) {
            const int itr = (y - y1) / NGEOPH;
            recpos[2 * itr] = x1 + ((x2 - x1) / (y2 - y1) * (y - y1));
            recpos[2 * itr + 1] = y;
        }
    }
    *ntr_out = ntr > 0 ? ntr : 0;
    return 0;
}

/* ---- gradient and model dumps (src/matrix_operations.c:24) ------------- */

int ifos_matrix_dump(const char *path_name, int NX, int NY, const float *m)
{
    size_t n = strlen(path_name);
    char *p = (char *)malloc(n + 5);
    if (!p) return ifos_set_error("out of host memory"), -1;
    memcpy(p, path_name, n);
    memcpy(p + n, ".bin", 5);
    const int rc = ifos_model_write(p, NX, NY, m);
    free(p);
    return rc;
}
