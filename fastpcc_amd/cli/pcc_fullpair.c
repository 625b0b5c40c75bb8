/*
 * pcc_fullpair — the correlation driver of PCC_fullpair_1b (src/PCC_fullpair_1b.c:
 * 144-502) on libfastpcc: the same command line, input formats (SAC file lists,
 * MSACS1), preprocessing (zero traces, clipping, outlier traces, whitening,
 * polarity), automatic pairing by start time, correlation types (PCC of power v,
 * WPCC, GNCC, 1-bit GNCC) and outputs (one SAC file per correlation, or one
 * ts-PWS binary file per channel pair).  The correlations and the preprocessing
 * run on the GPU through the library's C entry points.
 *
 *   pcc_fullpair filelist1 filelist2 [tl1= tl2= | nl1= nl2=] [pcc [v=]] [wpcc]
 *                [ccgn] [cc1b] [clip] [std=] [awhite=f1,f2] [isac|imsacs]
 *                [osac|obin[=prefix]] [Nmax=] [mincc=] [mindist=] [maxdist=]
 *                [pmin=] [pmax=] [VR=] [V=] [type=] [w0=] [NoAutoPairing] [verbose=]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fastpcc.h"
#include "fastpcc_io.h"

#define PI 3.14159265358979323846

typedef struct {
    double tl1, tl2, v;
    int nl1, nl2;
    int iformat, oformat; /* 1: SAC, 2: MSACS1 / binary */
    int pcc, wpcc, ccgn, cc1b, clip;
    int mincc, Nmax;
    double std, mindist, maxdist, pmin, pmax, VR;
    unsigned V;
    int type;
    double op1, awhite[2];
    const char *fin1, *fin2, *obinprefix;
    int autopair, acc, verbose;
} options;

typedef struct {
    fpcc_trace_header *hdr;
    float *x; /* [tr][N] */
    unsigned tr, N;
    float dt;
} station;

static int rd_double(double *v, const char *s)
{
    char *e;
    *v = strtod(s, &e);
    return e == s;
}

static int rd_int(int *v, const char *s)
{
    char *e;
    *v = (int)strtol(s, &e, 10);
    return e == s;
}

static int rd_pair(double *v, const char *s)
{
    char *e;
    v[0] = strtod(s, &e);
    if (e == s || *e != ',') return 1;
    const char *s2 = e + 1;
    v[1] = strtod(s2, &e);
    return e == s2;
}

static int parse(int argc, char **argv, options *o)
{
    int er = 0;
    for (int i = 3; i < argc; i++) {
        const char *a = argv[i];
        unsigned uv;
        if (!strncmp(a, "tl1=", 4)) er += rd_double(&o->tl1, a + 4);
        else if (!strncmp(a, "tl2=", 4)) er += rd_double(&o->tl2, a + 4);
        else if (!strncmp(a, "v=", 2)) er += rd_double(&o->v, a + 2);
        else if (!strncmp(a, "nl1=", 4)) er += rd_int(&o->nl1, a + 4);
        else if (!strncmp(a, "nl2=", 4)) er += rd_int(&o->nl2, a + 4);
        else if (!strncmp(a, "isac", 4)) o->iformat = 1;
        else if (!strncmp(a, "imsacs", 6)) o->iformat = 2;
        else if (!strncmp(a, "osac", 4)) o->oformat = 1;
        else if (!strncmp(a, "obin", 4)) {
            o->oformat = 2;
            if (!strncmp(a, "obin=", 5)) o->obinprefix = a + 5;
        } else if (!strncmp(a, "pcc", 3)) o->pcc = 1;
        else if (!strncmp(a, "wpcc", 4)) o->wpcc = 1;
        else if (!strncmp(a, "ccgn", 4)) o->ccgn = 1;
        else if (!strncmp(a, "cc1b", 4)) o->cc1b = 1;
        else if (!strncmp(a, "clip", 4)) o->clip = 1;
        else if (!strncmp(a, "std=", 4)) er += rd_double(&o->std, a + 4);
        else if (!strncmp(a, "mindist=", 8)) er += rd_double(&o->mindist, a + 8);
        else if (!strncmp(a, "maxdist=", 8)) er += rd_double(&o->maxdist, a + 8);
        else if (!strncmp(a, "pmin=", 5)) er += rd_double(&o->pmin, a + 5);
        else if (!strncmp(a, "pmax=", 5)) er += rd_double(&o->pmax, a + 5);
        else if (!strncmp(a, "VR=", 3)) er += rd_double(&o->VR, a + 3);
        else if (!strncmp(a, "mincc=", 6)) er += rd_int(&o->mincc, a + 6);
        else if (!strncmp(a, "Nmax=", 5)) er += rd_int(&o->Nmax, a + 5);
        else if (!strncmp(a, "V=", 2)) {
            int t = 0;
            er += rd_int(&t, a + 2);
            uv = (unsigned)t;
            o->V = uv;
        } else if (!strncmp(a, "type=", 5)) er += rd_int(&o->type, a + 5);
        else if (!strncmp(a, "w0=", 3)) er += rd_double(&o->op1, a + 3);
        else if (!strncmp(a, "awhite=", 7)) er += rd_pair(o->awhite, a + 7);
        else if (!strncmp(a, "NoAutoPairing", 13)) o->autopair = 0;
        else if (!strncmp(a, "verbose=", 8)) er += rd_int(&o->verbose, a + 8);
    }
    return er;
}

static int read_station(const char *fin, const options *o, station *st)
{
    memset(st, 0, sizeof *st);
    unsigned n = 0, N = (unsigned)o->Nmax;
    int r;
    if (o->iformat == 2) {
        if ((r = fpcc_msacs_read(fin, 0, NULL, NULL, &n, &N))) return r;
        st->hdr = calloc(n ? n : 1, sizeof(fpcc_trace_header));
        st->x = calloc((size_t)(n ? n : 1) * N, sizeof(float));
        if (!st->hdr || !st->x) return 4;
        if ((r = fpcc_msacs_read(fin, n, st->hdr, st->x, &n, &N))) return r;
        st->dt = n ? st->hdr[0].dt : 0.0f;  /* the file's npts (Nmax does not apply, as in the reference) */
    } else {
        if ((r = fpcc_sac_read_list(fin, &N, 0, NULL, NULL, &n, &st->dt))) return r;
        st->hdr = calloc(n ? n : 1, sizeof(fpcc_trace_header));
        st->x = calloc((size_t)(n ? n : 1) * N, sizeof(float));
        if (!st->hdr || !st->x) return 4;
        if ((r = fpcc_sac_read_list(fin, &N, n, st->hdr, st->x, &n, &st->dt))) return r;
    }
    st->tr = n;
    st->N = N;
    return 0;
}

/* Station location from the first trace: 0, or -4 when it is undefined (other
 * failures leave the location unknown too).  An MSACS1 file with a single trace
 * leaves lat/lon unchanged and returns 0, as ReadLocation_ManySacsFile does. */
static int read_location(const char *fin, const options *o, double *lat, double *lon)
{
    fpcc_trace_header h;
    if (o->iformat == 2) {
        unsigned n = 0, N = 0;
        if (fpcc_msacs_read(fin, 0, NULL, NULL, &n, &N)) return -2;
        if (n < 2) return 0;
        fpcc_trace_header *hs = calloc(n, sizeof *hs);
        float *x = calloc((size_t)n * N, sizeof(float));
        int r = (hs && x) ? fpcc_msacs_read(fin, n, hs, x, &n, &N) : -1;
        if (!r) h = hs[0];
        free(hs);
        free(x);
        if (r) return -4;
    } else {
        FILE *f = fopen(fin, "r");
        char line[1024];
        if (!f) return -2;
        char *s = fgets(line, sizeof line, f);
        fclose(f);
        if (!s) return -3;
        if ((s = strchr(line, '\n'))) *s = '\0';
        if (fpcc_sac_read(line, 0, &h, NULL)) return -4;
    }
    if (h.nostloc) return -4;
    *lat = h.stla;
    *lon = h.stlo;
    return 0;
}

/* Per-station preprocessing on the GPU over the station's traces (S = 1, W =
 * traces): signs, zero traces, rms, clipping, outliers; dropped traces are
 * removed from the list.  white != 0: whitening only (the traces are kept). */
static int preprocess_station(station *st, const options *o, int white)
{
    if (st->tr == 0) return 0;
    fpcc_job job;
    memset(&job, 0, sizeof job);
    job.n_stations = 1;
    job.n_windows = st->tr;
    job.win_len = st->N;
    job.station_stride = (unsigned long long)st->tr * st->N;
    fpcc_preproc pp;
    memset(&pp, 0, sizeof pp);
    if (white) {
        pp.band[0] = o->awhite[0];
        pp.band[1] = o->awhite[1];
        pp.dt = st->dt;
    } else {
        pp.clip = o->clip;
        pp.std_max = (float)o->std;
    }
    unsigned char *valid = malloc(st->tr);
    if (!valid) return 4;
    int r = fpcc_preprocess(&job, &pp, st->x, NULL, valid, NULL);
    if (!r && !white) {
        unsigned k = 0;
        for (unsigned t = 0; t < st->tr; t++) {
            if (!valid[t]) {
                if (o->verbose) printf("pcc_fullpair: skipping %s.%s.%s.%s, sequence %u (zero or outlier)\n",
                                       st->hdr[t].net, st->hdr[t].sta, st->hdr[t].loc, st->hdr[t].chn, t);
                continue;
            }
            if (k != t) {
                st->hdr[k] = st->hdr[t];
                memmove(st->x + (size_t)k * st->N, st->x + (size_t)t * st->N, sizeof(float) * st->N);
            }
            k++;
        }
        st->tr = k;
    }
    free(valid);
    return r;
}

/* Keeps the traces perm[0..n) in that order. */
static int reorder(station *st, const unsigned *perm, unsigned n)
{
    fpcc_trace_header *h = calloc(n ? n : 1, sizeof *h);
    float *x = calloc((size_t)(n ? n : 1) * st->N, sizeof(float));
    if (!h || !x) {
        free(h);
        free(x);
        return 4;
    }
    for (unsigned k = 0; k < n; k++) {
        h[k] = st->hdr[perm[k]];
        memcpy(x + (size_t)k * st->N, st->x + (size_t)perm[k] * st->N, sizeof(float) * st->N);
    }
    free(st->hdr);
    free(st->x);
    st->hdr = h;
    st->x = x;
    st->tr = n;
    return 0;
}

static int sort_station(station *st)
{
    unsigned *perm = malloc(sizeof(unsigned) * (st->tr ? st->tr : 1));
    if (!perm) return 4;
    fpcc_sort_traces(st->hdr, st->tr, perm);
    int r = reorder(st, perm, st->tr);
    free(perm);
    return r;
}

/* One SAC file per correlation (StoreInManySacs, src/PCC_fullpair_1b.c:680-705). */
static void store_sacs(float **y, unsigned L, unsigned Tr, int lag1, const station *s1, const station *s2, float dt,
                       const char *name)
{
    char dat[30], out[100];
    for (unsigned t = 0; t < Tr; t++) {
        const fpcc_trace_header *a = &s1->hdr[t], *b = &s2->hdr[t];
        snprintf(dat, sizeof dat, "_%s_%04d.%03d.%02d.%02d.%02d", name, a->year, a->yday, a->hour, a->min, a->sec);
        snprintf(out, sizeof out, "%s.%s.%s.%s.%s.%s.%s.%s%s.sac", a->net, a->sta, a->loc, a->chn, b->net, b->sta,
                 b->loc, b->chn, dat);
        if (fpcc_sac_write_cc(out, name, lag1 * dt, dt, y[t], (int)L, a, b))
            printf("wrsac: Error writing %s file\n", out);
    }
}

/* One binary file per channel pair (StoreInManyBins, src/PCC_fullpair_1b.c:774-829). */
static void store_bins(const float *ybuf, unsigned L, unsigned Tr, int lag1, const station *s1, const station *s2,
                       float dt, const char *name, const char *prefix)
{
    unsigned *set = calloc(Tr ? Tr : 1, sizeof(unsigned));
    char *done = calloc(Tr + 1, 1);
    char fname[80];
    if (!set || !done) {
        free(set);
        free(done);
        return;
    }
    unsigned next = 0;
    while (next < Tr) {
        const char *c1 = s1->hdr[next].chn, *c2 = s2->hdr[next].chn;
        unsigned n = 0;
        for (unsigned t = 0; t < Tr; t++)
            if (!strcmp(s1->hdr[t].chn, c1) && !strcmp(s2->hdr[t].chn, c2)) {
                done[t] = 1;
                set[n++] = t;
            }
        const fpcc_trace_header *a = &s1->hdr[next], *b = &s2->hdr[next];
        if (prefix)
            snprintf(fname, sizeof fname, "%s.%s.%s%c.%s.%s%c_%s.bin", prefix, a->sta, a->loc, c1[2], b->sta, b->loc,
                     c2[2], name);
        else
            snprintf(fname, sizeof fname, "%s.%s%c.%s.%s%c_%s.bin", a->sta, a->loc, c1[2], b->sta, b->loc, c2[2],
                     name);
        if (fpcc_obin_write(fname, name, ybuf, L, lag1, dt, s1->hdr, s2->hdr, set, n))
            printf("StoreInBin: cannot create %s file\n", fname);
        while (done[next] != 0) next++;
    }
    free(set);
    free(done);
}

static int run(options *o)
{
    station s1, s2;
    double lat1 = -90, lon1 = 0, lat2 = 90, lon2 = 0, gcarc = 0;
    int stloc = 1, r;

    if (o->iformat != 1 && o->iformat != 2) {
        printf("pcc_fullpair: Unknown format.");
        return 5;
    }
    if (read_location(o->fin1, o, &lat1, &lon1)) stloc = 0;
    if (!o->acc && read_location(o->fin2, o, &lat2, &lon2)) stloc = 0;
    if (o->acc) {
        lat2 = lat1;
        lon2 = lon1;
    }
    if (!stloc) {
        if (o->mindist > 0 || o->maxdist > 0 || o->VR > 0)
            printf("pcc_fullpair: Warning, the station location is not available, following without mindist, "
                   "maxdist and VR.\n");
        o->mindist = o->maxdist = o->VR = 0;
    } else {
        const double p1 = lat1 * PI / 180., l1 = lon1 * PI / 180., p2 = lat2 * PI / 180., l2 = lon2 * PI / 180.;
        const double c = sin(p2) * sin(p1) + cos(p2) * cos(p1) * cos(l1 - l2);
        gcarc = (c > 1. ? 0. : acos(c)) * 180. / PI;
        if (o->mindist > 0 && gcarc < o->mindist) {
            printf("pcc_fullpair: The interstation distance is smaller than %f degrees on %s - %s.\n", o->mindist,
                   o->fin1, o->fin2);
            return 0;
        }
        if (o->maxdist > 0 && gcarc > o->maxdist) {
            printf("pcc_fullpair: The interstation distance is higher than %f degrees on %s - %s.\n", o->maxdist,
                   o->fin1, o->fin2);
            return 0;
        }
    }

    if ((r = read_station(o->fin1, o, &s1))) {
        printf("pcc_fullpair: Something went wrong when reading the data from station 1! (nerr = %d)\n", r);
        return r;
    }
    if (s1.tr == 0) {
        printf("pcc_fullpair: %s is empty.", o->fin1);
        return 0;
    }
    if ((int)s1.tr < o->mincc) {
        printf("pcc_fullpair: Too few sequences from station %s (%u < %d)\n", o->fin1, s1.tr, o->mincc);
        return 0;
    }
    if (o->acc) {
        s2 = s1;
    } else {
        if ((r = read_station(o->fin2, o, &s2))) {
            printf("pcc_fullpair: Something went wrong when reading the data from station 2! (nerr = %d)\n", r);
            return r;
        }
        if (s2.tr == 0) {
            printf("pcc_fullpair: %s is empty.", o->fin2);
            return 0;
        }
        if ((int)s2.tr < o->mincc) {
            printf("pcc_fullpair: Too few sequences from station %s (%u < %d)\n", o->fin2, s2.tr, o->mincc);
            return 0;
        }
        if (s1.N != s2.N) printf("pcc_fullpair: These traces have different lengths: %u:%u\n", s1.N, s2.N);
        if (s1.dt != s2.dt)
            printf("pcc_fullpair: These stations have different samplings: %11.9f:%11.9f\n", s1.dt, s2.dt);
    }
    const unsigned N = s1.N;
    const float dt = s2.dt;

    /* zero traces, rms, clipping, outliers (GPU) */
    if ((r = preprocess_station(&s1, o, 0))) return r;
    if (o->acc) s2 = s1;
    else if ((r = preprocess_station(&s2, o, 0))) return r;

    if (o->autopair) {
        if ((r = sort_station(&s1))) return r;
        if (o->acc) s2 = s1;
        else {
            if ((r = sort_station(&s2))) return r;
            unsigned *i1 = malloc(sizeof(unsigned) * (s1.tr ? s1.tr : 1));
            unsigned *i2 = malloc(sizeof(unsigned) * (s2.tr ? s2.tr : 1));
            unsigned np = 0;
            if (!i1 || !i2) return 4;
            fpcc_pair_traces(s1.hdr, s1.tr, s2.hdr, s2.tr, i1, i2, &np);
            if ((r = reorder(&s1, i1, np)) || (r = reorder(&s2, i2, np))) return r;
            free(i1);
            free(i2);
        }
    } else if (!o->acc) {
        r = fpcc_check_pairs(s1.hdr, s1.tr, s2.hdr, s2.tr);
        if (r) {
            printf("pcc_fullpair: Something went wrong when checking the pairs! (nerr = %d)\n", r);
            return r;
        }
    }
    const unsigned Tr = s1.tr;

    /* whitening of each station's paired traces (GPU) */
    if (o->awhite[0] > 0 && o->awhite[1] > o->awhite[0]) {
        if ((r = preprocess_station(&s1, o, 1))) return r;
        if (o->acc) s2 = s1;
        else if ((r = preprocess_station(&s2, o, 1))) return r;
    }

    int lag1 = o->nl1 ? o->nl1 : (o->tl1 ? (int)round(o->tl1 / (double)dt) : 0);
    int lag2 = o->nl2 ? o->nl2 : (o->tl2 ? (int)round(o->tl2 / (double)dt) : 0);
    if (lag1 > lag2) {
        int t = lag1;
        lag1 = lag2;
        lag2 = t;
    }
    if ((unsigned)abs(lag1) >= N || (unsigned)abs(lag2) >= N) {
        printf("pcc_fullpair: TOO LARGE LAGS!!! The modulus of the Lags have to be lower than the sequence length.\n");
        return 6;
    }
    const unsigned L = (unsigned)(lag2 - lag1 + 1);
    printf("Lag1 = %d smpl (%d s), Lag2 = %d smpl (%d s), ", lag1, (int)(lag1 * dt), lag2, (int)(lag2 * dt));
    printf("L = %u smpl (%d s), N = %u smpl, Tr = %u, gcarc = %f deg\n", L, (int)((L - 1) * dt), N, Tr, gcarc);
    if ((int)Tr <= o->mincc) {
        if (!Tr) printf("NO INTERSTATION CORRELATION TO BE COMPUTED.\n");
        else printf("ONLY %u INTERSTATION CORRELATION COULD BE COMPUTED.\n", Tr);
        return 0;
    }

    double pmin = 1.25 * sqrt(2) * PI, pmax = pmin * pow(2, 3.5);
    if (o->wpcc && o->pmin != 0 && o->pmax != 0) {
        pmin = o->pmin / dt;
        if (o->VR == 0) pmax = o->pmax / dt;
        else {
            pmax = 111.19 * gcarc / (3 * o->VR);
            if (pmax > o->pmax) pmax = o->pmax;
            pmax /= dt;
        }
        printf("pmin = %f smpl (%f s), pmax = %f smpl (%f s)\n", pmin, pmin * dt, pmax, pmax * dt);
    }
    char nick[16];
    if (o->v == 2) strcpy(nick, "pcc2");
    else if (o->v == 1) strcpy(nick, "pcc1");
    else snprintf(nick, sizeof nick, "pcc%.1f", o->v);

    float *ybuf = calloc((size_t)Tr * L, sizeof(float));
    float **y = malloc(sizeof(float *) * Tr), **x1 = malloc(sizeof(float *) * Tr), **x2 = malloc(sizeof(float *) * Tr);
    if (!ybuf || !y || !x1 || !x2) {
        printf("pcc_fullpair: Out of memory on the output array (%u x %u)\n", L, Tr);
        return 4;
    }
    for (unsigned t = 0; t < Tr; t++) {
        y[t] = ybuf + (size_t)t * L;
        /* CorrectRevesedPolarity (src/rotlib.c:14-34) on the paired traces */
        x1[t] = s1.x + (size_t)t * N;
        x2[t] = s2.x + (size_t)t * N;
    }
    for (unsigned t = 0; t < Tr; t++)
        if (fpcc_polarity(&s1.hdr[t]) < 0)
            for (unsigned n = 0; n < N; n++) x1[t][n] = -x1[t][n];
    if (!o->acc)
        for (unsigned t = 0; t < Tr; t++)
            if (fpcc_polarity(&s2.hdr[t]) < 0)
                for (unsigned n = 0; n < N; n++) x2[t][n] = -x2[t][n];

    struct {
        int on;
        const char *name;
        int kind;
    } todo[4] = {{o->pcc, nick, 0}, {o->wpcc, "wpcc2", 1}, {o->ccgn, "ccgn", 2}, {o->cc1b, "cc1b", 3}};
    for (int k = 0; k < 4; k++) {
        if (!todo[k].on) continue;
        memset(ybuf, 0, sizeof(float) * (size_t)Tr * L);
        int e = 0;
        switch (todo[k].kind) {
        case 0:
            if (o->v == 2) e = pcc2_set(y, x1, x2, N, Tr, lag1, lag2);
            else if (o->v == 1) e = pcc1_set(y, x1, x2, (int)N, Tr, lag1, lag2);
            else e = pcc_set(y, x1, x2, (int)N, Tr, o->v, lag1, lag2);
            break;
        case 1: e = tspcc2_set(y, x1, x2, N, Tr, lag1, lag2, pmin, pmax, o->V, o->type, o->op1); break;
        case 2: e = ccgn_set(y, x1, x2, N, Tr, lag1, lag2); break;
        default: e = cc1b_set(y, x1, x2, N, Tr, lag1, lag2); break;
        }
        if (e) printf("pcc_fullpair: %s returned %d\n", todo[k].name, e);
        if (o->oformat == 1) store_sacs(y, L, Tr, lag1, &s1, &s2, dt, todo[k].name);
        else if (o->oformat == 2) store_bins(ybuf, L, Tr, lag1, &s1, &s2, dt, todo[k].name, o->obinprefix);
    }
    free(ybuf);
    free(y);
    free(x1);
    free(x2);
    return 0;
}

int main(int argc, char **argv)
{
    options o;
    memset(&o, 0, sizeof o);
    o.v = 2;
    o.iformat = 1;
    o.oformat = 1;
    o.V = 2;
    o.type = -3;
    o.autopair = 1;
    o.verbose = 1;
    if (argc < 3) {
        puts("USAGE: pcc_fullpair filelist1 filelist2 parameters  (the parameters of PCC_fullpair_1b)");
        return 0;
    }
    o.fin1 = argv[1];
    o.fin2 = argv[2];
    o.acc = strcmp(o.fin1, o.fin2) == 0;
    int er = parse(argc, argv, &o);
    if (!o.autopair && o.std > 0) {
        printf("pcc_fullpair: Warning, the std option cannot be used without automatic trace pairing.\n");
        o.std = 0;
    }
    if (o.pmin > o.pmax) {
        printf("pcc_fullpair: Error, pmax has to be higher than pmin.\n");
        er = 1;
    }
    return er ? er : run(&o);
}
