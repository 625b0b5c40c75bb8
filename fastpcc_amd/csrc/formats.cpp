// On-disk formats and trace pairing of the correlation driver (host code).
//
// SAC binary layout (version 6): 70 floats, 40 ints, 24 strings of 8 bytes (the
// second, kevnm, of 16) = 632 header bytes, then npts float samples.  Undefined
// values are -12345 (floats and ints) and "-12345  " (strings).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/fastpcc_io.h"

static_assert(sizeof(fpcc_trace_header) == 120, "fpcc_trace_header must keep t_HeaderInfo's layout");

namespace {

constexpr int kSacHeaderBytes = 632;
constexpr float kUndefF = -12345.0f;
constexpr int32_t kUndefI = -12345;

// float indices
enum { F_DELTA = 0, F_DEPMIN = 1, F_DEPMAX = 2, F_B = 5, F_E = 6, F_O = 7, F_STLA = 31, F_STLO = 32, F_STEL = 33,
       F_STDP = 34, F_EVLA = 35, F_EVLO = 36, F_EVEL = 37, F_EVDP = 38, F_DEPMEN = 56, F_CMPAZ = 57,
       F_CMPINC = 58 };
// int indices
enum { I_NZYEAR = 0, I_NZJDAY = 1, I_NZHOUR = 2, I_NZMIN = 3, I_NZSEC = 4, I_NZMSEC = 5, I_NVHDR = 6, I_NPTS = 9,
       I_IFTYPE = 15, I_IZTYPE = 17, I_LEVEN = 35, I_LPSPOL = 36, I_LOVROK = 37, I_LCALDA = 38 };
// string words (8 bytes each; kevnm takes words 1 and 2)
enum { K_KSTNM = 0, K_KEVNM = 1, K_KHOLE = 3, K_KO = 4, K_KUSER0 = 17, K_KUSER1 = 18, K_KUSER2 = 19, K_KCMPNM = 20,
       K_KNETWK = 21, K_KINST = 23 };
constexpr int kItime = 1, kIo = 11;

struct SacHeader {
    float f[70];
    int32_t i[40];
    char k[192];

    char *str(int word) { return k + word * 8; }
    const char *str(int word) const { return k + word * 8; }
};
static_assert(sizeof(SacHeader) == kSacHeaderBytes, "SAC header is 632 bytes");

uint32_t bswap32(uint32_t v) { return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24); }

void swap_words(void *p, size_t nwords)
{
    uint32_t *w = (uint32_t *)p;
    for (size_t n = 0; n < nwords; ++n) w[n] = bswap32(w[n]);
}

bool is_leap(int year) { return (year % 4) == 0 && ((year % 100) != 0 || (year % 400) == 0); }

// Seconds since 1970-01-01 of day yday (1-based) of year at hour:min:sec, as
// my_timegm computes it for tm_mon = 0, tm_mday = yday.
int64_t start_time(int year, int yday, int hour, int min, int sec)
{
    int64_t days = 0;
    for (int y = 1970; y < year; ++y) days += is_leap(y) ? 366 : 365;
    days += yday - 1;
    return ((days * 24 + hour) * 60 + min) * 60 + sec;
}

// A header string as the reference keeps it: 8 characters, cut at the first
// blank, empty when undefined.
void get_string(const char *src, char *dst)
{
    memcpy(dst, src, 8);
    dst[8] = '\0';
    if (!strncmp(dst, "-12345", 6)) {
        dst[0] = '\0';
        return;
    }
    if (char *sp = (char *)memchr(dst, ' ', 8)) *sp = '\0';
}

// A header string as sacio sets it: up to len characters, blank-padded.
void set_string(char *dst, int len, const char *src)
{
    const size_t n = std::min(strlen(src), (size_t)len);
    memset(dst, ' ', len);
    memcpy(dst, src, n);
}

int read_header(FILE *f, SacHeader &h, bool &swapped)
{
    if (fread(&h, sizeof h, 1, f) != 1) return -2;
    swapped = false;
    if (h.i[I_NVHDR] != 6 && h.i[I_NVHDR] != 7) {
        const uint32_t v = bswap32((uint32_t)h.i[I_NVHDR]);
        if (v != 6 && v != 7) return -3;
        swap_words(h.f, 70);
        swap_words(h.i, 40);
        swapped = true;
    }
    return 0;
}

void header_info(const SacHeader &h, fpcc_trace_header *o)
{
    memset(o, 0, sizeof *o);
    o->npts = h.i[I_NPTS];
    o->dt = h.f[F_DELTA];
    o->b = h.f[F_B];
    o->year = h.i[I_NZYEAR];
    o->yday = h.i[I_NZJDAY];
    o->hour = h.i[I_NZHOUR];
    o->min = h.i[I_NZMIN];
    o->sec = h.i[I_NZSEC];
    o->msec = h.i[I_NZMSEC];
    get_string(h.str(K_KNETWK), o->net);
    get_string(h.str(K_KSTNM), o->sta);
    get_string(h.str(K_KCMPNM), o->chn);
    get_string(h.str(K_KHOLE), o->loc);
    o->stla = h.f[F_STLA];
    o->stlo = h.f[F_STLO];
    if (o->stla == kUndefF || o->stlo == kUndefF) o->nostloc = 1;
    o->stel = h.f[F_STEL] == kUndefF ? 0.0f : h.f[F_STEL];
    o->stdp = h.f[F_STDP] == kUndefF ? 0.0f : h.f[F_STDP];
    o->cmpaz = h.f[F_CMPAZ];
    o->cmpinc = h.f[F_CMPINC];
    if (o->cmpaz == kUndefF || o->cmpinc == kUndefF) o->nocmp = 1;
    o->t = start_time(o->year, o->yday, o->hour, o->min, o->sec);
}

std::vector<std::string> read_list(const char *filelist, bool &ok)
{
    std::vector<std::string> out;
    ok = false;
    FILE *f = fopen(filelist, "r");
    if (!f) return out;
    char line[1024];
    while (fgets(line, sizeof line, f)) {
        if (char *nl = strchr(line, '\n')) *nl = '\0';
        out.emplace_back(line);
    }
    fclose(f);
    ok = true;
    return out;
}

double time_diff(const fpcc_trace_header &a, const fpcc_trace_header &b)
{
    return difftime((time_t)a.t, (time_t)b.t) + (double)(a.msec - b.msec) / 1000;
}

}  // namespace

extern "C" int fpcc_sac_read(const char *path, unsigned nmax, fpcc_trace_header *hdr, float *data)
{
    if (!path || !hdr) return -1;
    FILE *f = fopen(path, "rb");
    if (!f) return -2;
    SacHeader h;
    bool swapped = false;
    int r = read_header(f, h, swapped);
    if (r) {
        fclose(f);
        return r;
    }
    header_info(h, hdr);
    if (data && h.i[I_NPTS] > 0) {
        const size_t n = nmax ? std::min<size_t>((size_t)h.i[I_NPTS], nmax) : (size_t)h.i[I_NPTS];
        if (fread(data, sizeof(float), n, f) != n) {
            fclose(f);
            return -2;
        }
        if (swapped) swap_words(data, n);
    }
    fclose(f);
    return 0;
}

extern "C" int fpcc_sac_read_list(const char *filelist, unsigned *N, unsigned max_traces, fpcc_trace_header *hdrs,
                                  float *x, unsigned *n_traces, float *dt)
{
    if (!filelist || !N || !n_traces) return -1;
    *n_traces = 0;
    bool ok = false;
    const std::vector<std::string> files = read_list(filelist, ok);
    if (!ok || files.empty()) return -2;
    fpcc_trace_header first;
    if (fpcc_sac_read(files[0].c_str(), 0, &first, nullptr)) return -2;
    const unsigned n = *N ? *N : (unsigned)std::max(first.npts, 0);
    const float dt1 = first.dt, b1 = first.b;
    std::vector<float> sig(x ? n : 0);
    unsigned kept = 0;
    for (const std::string &name : files) {
        fpcc_trace_header h;
        if (fpcc_sac_read(name.c_str(), n, &h, x ? sig.data() : nullptr)) return -2;
        if ((unsigned)std::max(h.npts, 0) < n) continue;  // too short
        if (h.npts > (int)n) h.npts = (int)n;             // cut
        if (std::fabs(h.dt - dt1) > dt1 * 0.001) continue;
        if (std::fabs(b1 - h.b) > dt1) continue;
        if (hdrs || x) {
            if (kept >= max_traces) return -5;
            if (hdrs) hdrs[kept] = h;
            if (x) memcpy(x + (size_t)kept * n, sig.data(), sizeof(float) * n);
        }
        ++kept;
    }
    *N = n;
    *n_traces = kept;
    if (dt) *dt = dt1;
    return 0;
}

extern "C" int fpcc_msacs_read(const char *path, unsigned max_traces, fpcc_trace_header *hdrs, float *x,
                               unsigned *n_traces, unsigned *N)
{
    if (!path || !n_traces || !N) return -1;
    FILE *f = fopen(path, "rb");
    if (!f) return -2;
    struct {
        char id[8];
        int32_t nseq, npts, single;
    } mh;
    if (fread(&mh, sizeof mh, 1, f) != 1 || memcmp(mh.id, "MSACS1", 7) || mh.npts <= 0 || mh.nseq < 0) {
        fclose(f);
        return -2;
    }
    *n_traces = (unsigned)mh.nseq;
    *N = (unsigned)mh.npts;
    if (!hdrs && !x) {
        fclose(f);
        return 0;
    }
    if ((unsigned)mh.nseq > max_traces) {
        fclose(f);
        return -5;
    }
    std::vector<float> skip;
    for (int tr = 0; tr < mh.nseq; ++tr) {
        fpcc_trace_header h;
        if (fread(&h, sizeof h, 1, f) != 1) {
            fclose(f);
            return -2;
        }
        const size_t np = std::min<size_t>((size_t)std::max(h.npts, 0), (size_t)mh.npts);
        float *dst = x ? x + (size_t)tr * mh.npts : nullptr;
        if (!dst) {
            skip.resize(np);
            dst = skip.data();
        }
        if (fread(dst, sizeof(float), np, f) != np) {
            fclose(f);
            return -2;
        }
        if (x && np < (size_t)mh.npts) memset(dst + np, 0, sizeof(float) * (mh.npts - np));
        if (hdrs) hdrs[tr] = h;
    }
    fclose(f);
    return 0;
}

extern "C" int fpcc_msacs_write(const char *path, const fpcc_trace_header *hdrs, const float *x, unsigned n_traces,
                                unsigned N)
{
    if (!path || (!hdrs && n_traces) || (!x && n_traces)) return -1;
    FILE *f = fopen(path, "wb");
    if (!f) return -2;
    struct {
        char id[8];
        int32_t nseq, npts, single;
    } mh;
    memset(&mh, 0, sizeof mh);
    strcpy(mh.id, "MSACS1");
    mh.nseq = (int32_t)n_traces;
    mh.npts = (int32_t)N;
    mh.single = 1;
    bool ok = fwrite(&mh, sizeof mh, 1, f) == 1;
    for (unsigned tr = 0; ok && tr < n_traces; ++tr)
        ok = fwrite(&hdrs[tr], sizeof(fpcc_trace_header), 1, f) == 1 &&
             fwrite(x + (size_t)tr * N, sizeof(float), N, f) == N;
    fclose(f);
    return ok ? 0 : -2;
}

extern "C" int fpcc_sac_write_cc(const char *path, const char *method, float beg, float dt, const float *y, int nsamp,
                                 const fpcc_trace_header *h1, const fpcc_trace_header *h2)
{
    if (!path || !method || !y || nsamp < 1 || !h1 || !h2) return -1;
    SacHeader h;
    for (float &v : h.f) v = kUndefF;
    for (int32_t &v : h.i) v = kUndefI;
    for (int w = 0; w < 24; ++w) memcpy(h.k + w * 8, "-12345  ", 8);
    h.i[I_NVHDR] = 6;
    h.i[I_IFTYPE] = kItime;
    h.i[I_LEVEN] = 1;
    h.i[I_LPSPOL] = 0;
    h.i[I_LOVROK] = 1;
    h.i[I_LCALDA] = 1;
    h.i[I_NPTS] = nsamp;
    h.f[F_DELTA] = dt;
    set_string(h.str(K_KINST), 8, method);
    if (!h1->nostloc && !h2->nostloc) {
        h.f[F_STLA] = h2->stla;
        h.f[F_STLO] = h2->stlo;
        h.f[F_STEL] = h2->stel;
        h.f[F_STDP] = h2->stdp;
        set_string(h.str(K_KNETWK), 8, h2->net);
        set_string(h.str(K_KSTNM), 8, h2->sta);
        set_string(h.str(K_KHOLE), 8, h2->loc);
        set_string(h.str(K_KCMPNM), 8, h2->chn);
        h.f[F_EVLA] = h1->stla;
        h.f[F_EVLO] = h1->stlo;
        h.f[F_EVEL] = h1->stel;
        h.f[F_EVDP] = h1->stdp;
        set_string(h.str(K_KUSER0), 8, h1->net);
        set_string(h.str(K_KEVNM), 16, h1->sta);
        set_string(h.str(K_KUSER1), 8, h1->loc);
        set_string(h.str(K_KUSER2), 8, h1->chn);
    }
    h.i[I_NZYEAR] = h1->year;
    h.i[I_NZJDAY] = h1->yday;
    h.i[I_NZHOUR] = h1->hour;
    h.i[I_NZMIN] = h1->min;
    h.i[I_NZSEC] = h1->sec;
    h.i[I_NZMSEC] = h1->msec;
    const float lag0 = (float)(time_diff(*h1, *h2) + (h1->b - h2->b));
    h.f[F_B] = beg;
    h.f[F_O] = lag0;
    set_string(h.str(K_KO), 8, "0LagTime");
    h.i[I_IZTYPE] = kIo;
    // quantities sacio's writer derives from the samples
    float mn = y[0], mx = y[0];
    double sum = 0;
    for (int n = 0; n < nsamp; ++n) {
        mn = std::min(mn, y[n]);
        mx = std::max(mx, y[n]);
        sum += y[n];
    }
    h.f[F_DEPMIN] = mn;
    h.f[F_DEPMAX] = mx;
    h.f[F_DEPMEN] = (float)(sum / nsamp);
    h.f[F_E] = beg + (nsamp - 1) * dt;
    FILE *f = fopen(path, "wb");
    if (!f) return -2;
    const bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(y, sizeof(float), nsamp, f) == (size_t)nsamp;
    fclose(f);
    return ok ? 0 : -2;
}

extern "C" int fpcc_obin_write(const char *path, const char *method, const float *y, unsigned L, int lag1, float dt,
                               const fpcc_trace_header *h1, const fpcc_trace_header *h2, const unsigned *set,
                               unsigned nset)
{
    if (!path || !method || !y || !h1 || !h2 || !set || nset == 0 || L == 0) return -1;
    struct {
        char method[8], net1[8], sta1[8], loc1[8], chn1[8], net2[8], sta2[8], loc2[8], chn2[8];
        float stlat1, stlon1, stel1, stlat2, stlon2, stel2;
        uint32_t nlags, nseq;
        float tlength, lag1, lag2;
    } cc;
    static_assert(sizeof cc == 116, "t_ccheader is 116 bytes");
    memset(&cc, 0, sizeof cc);
    const fpcc_trace_header &a = h1[set[0]], &b = h2[set[0]];
    memcpy(cc.method, method, strnlen(method, 8));
    memcpy(cc.net1, a.net, 8);
    memcpy(cc.sta1, a.sta, 8);
    memcpy(cc.loc1, a.loc, 8);
    memcpy(cc.chn1, a.chn, 8);
    cc.stlat1 = a.stla;
    cc.stlon1 = a.stlo;
    cc.stel1 = a.stel;
    memcpy(cc.net2, b.net, 8);
    memcpy(cc.sta2, b.sta, 8);
    memcpy(cc.loc2, b.loc, 8);
    memcpy(cc.chn2, b.chn, 8);
    cc.stlat2 = b.stla;
    cc.stlon2 = b.stlo;
    cc.stel2 = b.stel;
    cc.nlags = L;
    cc.nseq = nset;
    cc.tlength = dt * (L - 1);
    cc.lag1 = dt * lag1;
    cc.lag2 = dt * (float)((unsigned)lag1 + L - 1);  // unsigned arithmetic, as the reference's int + unsigned
    std::vector<int64_t> times(nset);
    std::vector<float> lag0(nset);
    for (unsigned k = 0; k < nset; ++k) times[k] = h1[set[k]].t;
    const fpcc_trace_header &last1 = h1[set[nset - 1]];
    const float l0 = (float)(time_diff(last1, b) + (last1.b - b.b));
    for (unsigned k = 0; k < nset; ++k) lag0[k] = l0;
    FILE *f = fopen(path, "wb");
    if (!f) return -2;
    bool ok = fwrite(&cc, sizeof cc, 1, f) == 1 && fwrite(times.data(), sizeof(int64_t), nset, f) == nset &&
              fwrite(lag0.data(), sizeof(float), nset, f) == nset;
    for (unsigned k = 0; ok && k < nset; ++k) ok = fwrite(y + (size_t)set[k] * L, sizeof(float), L, f) == L;
    fclose(f);
    return ok ? 0 : -2;
}

extern "C" int fpcc_sort_traces(const fpcc_trace_header *hdrs, unsigned n, unsigned *perm)
{
    if ((!hdrs || !perm) && n) return -1;
    for (unsigned k = 0; k < n; ++k) perm[k] = k;
    // the reference's comparator: the truncated difference in seconds
    std::stable_sort(perm, perm + n, [&](unsigned a, unsigned b) { return (int)time_diff(hdrs[a], hdrs[b]) < 0; });
    return 0;
}

extern "C" int fpcc_pair_traces(const fpcc_trace_header *h1, unsigned n1, const fpcc_trace_header *h2, unsigned n2,
                                unsigned *idx1, unsigned *idx2, unsigned *n_pairs)
{
    if (!n_pairs || ((!h1 || !idx1) && n1) || ((!h2 || !idx2) && n2)) return -1;
    unsigned a = 0, b = 0, k = 0;
    while (a < n1 && b < n2) {
        const double td = time_diff(h1[a], h2[b]);
        if (std::fabs(td) < 0.5) {
            idx1[k] = a++;
            idx2[k] = b++;
            ++k;
        } else if (td < 0) {
            ++a;
        } else {
            ++b;
        }
    }
    *n_pairs = k;
    return 0;
}

extern "C" int fpcc_check_pairs(const fpcc_trace_header *h1, unsigned n1, const fpcc_trace_header *h2, unsigned n2)
{
    if (!h1 || !h2) return -1;
    if (n1 != n2) return -2;
    int bad = 0;
    for (unsigned k = 0; k < n1; ++k)
        if (std::fabs(time_diff(h1[k], h2[k])) > 0.5) ++bad;
    return bad;
}

extern "C" int fpcc_polarity(fpcc_trace_header *hdr)
{
    if (!hdr) return 1;
    if (hdr->cmpinc == 180) {
        hdr->cmpinc = 0;
        return -1;
    }
    if (hdr->cmpinc == 90 && hdr->cmpaz < 315 && hdr->cmpaz > 135) {
        hdr->cmpaz -= 180;
        if (hdr->cmpaz < 0) hdr->cmpaz += 360;
        return -1;
    }
    return 1;
}
