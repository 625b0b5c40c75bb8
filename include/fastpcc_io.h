/*
 * fastpcc_io.h — on-disk formats and trace pairing of the correlation driver.
 *
 * The reference reads SAC files through the external sacio library and its own
 * MSACS1 container, pairs the traces of two stations by their start times, and
 * writes correlations as SAC files or one binary file per channel pair for the
 * ts-PWS stack.  These entry points do the same on the host (plain C++ in
 * libfastpcc.so, no GPU), so a C driver can feed the correlation entry points of
 * fastpcc.h from the same files the reference reads and write the same files it
 * writes.  Formats are read and written little-endian (big-endian SAC files are
 * detected by their header version and swapped on reading).
 */
#ifndef FASTPCC_IO_H
#define FASTPCC_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One trace's metadata: t_HeaderInfo of src/ReadManySacs.h:7-30 with the same
 * field order, types and size (120 bytes on x86-64), so MSACS1 files written by
 * either program read in the other.  t is the start time in seconds since
 * 1970-01-01 from year/yday/hour/min/sec (my_timegm, src/ReadManySacs.c:62-84);
 * nostloc / nocmp are 1 when stla/stlo or cmpaz/cmpinc are undefined. */
typedef struct fpcc_trace_header {
    int32_t npts;
    float dt;
    float b;
    int32_t year, yday, hour, min, sec, msec;
    int64_t t;
    float stla, stlo, stel, stdp;
    float cmpaz, cmpinc;
    char sta[9], net[9], chn[9], loc[9];
    int32_t nostloc, nocmp;
} fpcc_trace_header;

/* One SAC file (the per-file part of ReadManySacs, src/ReadManySacs.c:190-245):
 * header fields as the reference extracts them (strings cut at the first blank,
 * undefined strings empty), and the first min(npts, nmax) samples into data when
 * data != NULL (nmax = 0: all).  Returns 0, -2 (cannot open / short file) or -3
 * (not a SAC file). */
int fpcc_sac_read(const char *path, unsigned nmax, fpcc_trace_header *hdr, float *data);

/* ReadManySacs (src/ReadManySacs.c:163-318): the SAC files named one per line in
 * filelist.  *N in: samples per trace (0: npts of the first file); out: the value
 * used.  Traces shorter than N, with a sampling interval off by more than 0.1% or
 * a begin time off by more than one sample from the first file's are skipped;
 * longer ones are cut to N.  hdrs[max_traces] and x[max_traces][N] receive the
 * kept traces (x may be NULL for headers only; both NULL to count).  *n_traces:
 * kept traces; *dt: the first file's sampling interval.  Returns 0, -2 (list or
 * file unreadable) or -5 (more kept traces than max_traces). */
int fpcc_sac_read_list(const char *filelist, unsigned *N, unsigned max_traces, fpcc_trace_header *hdrs, float *x,
                       unsigned *n_traces, float *dt);

/* MSACS1 container (Read_ManySacsFile / Write_ManySacsFile,
 * src/ReadManySacs.c:561-646): {"MSACS1", nseq, npts, 1} then per trace its
 * header and npts samples.  Reading: *n_traces and *N from the file; hdrs/x as
 * above (NULL to query).  Returns 0, -2 (unreadable or not MSACS1) or -5. */
int fpcc_msacs_read(const char *path, unsigned max_traces, fpcc_trace_header *hdrs, float *x, unsigned *n_traces,
                    unsigned *N);
int fpcc_msacs_write(const char *path, const fpcc_trace_header *hdrs, const float *x, unsigned n_traces, unsigned N);

/* One correlation as a SAC file (wrsac, src/PCC_fullpair_1b.c:831-879): npts,
 * delta, b = beg, kinst = method; station 2 as the station (stla, stlo, stel,
 * stdp, knetwk, kstnm, khole, kcmpnm) and station 1 as the event (evla, evlo,
 * evel, evdp, kuser0, kevnm, kuser1, kuser2) when both locations are known; the
 * reference time of trace 1 (nz*), o = the start-time difference of the two
 * traces, ko = "0LagTime", iztype = IO, lcalda = true.  Returns 0 or -2. */
int fpcc_sac_write_cc(const char *path, const char *method, float beg, float dt, const float *y, int nsamp,
                      const fpcc_trace_header *h1, const fpcc_trace_header *h2);

/* Correlations of one channel pair in the ts-PWS binary format (StoreInBin,
 * src/PCC_fullpair_1b.c:707-772, header t_ccheader of src/sac2bin.h): the 116-
 * byte header filled from the pair (set[0]) of trace headers, then nset start
 * times (int64, trace 1), nset lag-0 times (float) and nset rows of L floats
 * (y + set[k] * L).  The lag-0 time of every row is computed from trace 1 of the
 * last pair and trace 2 of the first pair, as the reference does.  Returns 0 or -2. */
int fpcc_obin_write(const char *path, const char *method, const float *y, unsigned L, int lag1, float dt,
                    const fpcc_trace_header *h1, const fpcc_trace_header *h2, const unsigned *set, unsigned nset);

/* SortTraces (src/PCC_fullpair_1b.c:927-986): perm[n] = the order of the traces
 * by start time (seconds, then milliseconds; stable). */
int fpcc_sort_traces(const fpcc_trace_header *hdrs, unsigned n, unsigned *perm);

/* MakePairedLists (src/PCC_fullpair_1b.c:988-1041): walks two time-sorted lists
 * and keeps the traces whose start times differ by less than 0.5 s (each trace in
 * at most one pair); idx1[k], idx2[k] are the kept positions, *n_pairs their
 * number. */
int fpcc_pair_traces(const fpcc_trace_header *h1, unsigned n1, const fpcc_trace_header *h2, unsigned n2,
                     unsigned *idx1, unsigned *idx2, unsigned *n_pairs);

/* CheckPairs (src/PCC_fullpair_1b.c:1043-1063): 0 when the lists pair line by
 * line within 0.5 s, -2 for different lengths, else the number of mismatches. */
int fpcc_check_pairs(const fpcc_trace_header *h1, unsigned n1, const fpcc_trace_header *h2, unsigned n2);

/* CorrectRevesedPolarity's rule (src/rotlib.c:14-34) for one trace: returns -1
 * (and updates cmpinc / cmpaz as the reference does) when the trace's samples are
 * to be negated, else +1. */
int fpcc_polarity(fpcc_trace_header *hdr);

#ifdef __cplusplus
}
#endif

#endif /* FASTPCC_IO_H */
