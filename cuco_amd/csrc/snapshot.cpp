// Snapshot files and the output-time schedule, host side.
//
// File format of savepositions (io.c:9-40) and read_ic (read_ic.c:9-45):
// one HEADER record (allvars.h:47-57, 32 bytes) followed by NumPart
// PARTICLE records (allvars.h:35-45, 60 bytes), native byte order, no
// padding or markers.  The output list is read_outputlist's whitespace
// separated list of scale factors (begrun.c:80-102), and the next output
// time on the integer timeline is find_next_outputtime's (run.c:75-104).
//
// Differences from the reference: failures return a status instead of
// exiting, a short or oversized file is rejected instead of read partially,
// and the list is read into doubles (the reference reads "%g" into float;
// the values are rounded to float here as well, so the integer output
// times agree).
//
// Checkpoint files (no counterpart in the reference, which cannot restart):
// their header is checked here, host only; the handle's state is written
// and restored by the engine.
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "cuco_engine.h"

namespace cuco {
void set_error(const char *fmt, ...);
}

using cuco::set_error;

static_assert(sizeof(cuco_header) == 32, "HEADER is 32 bytes");
static_assert(sizeof(cuco_particle) == 60, "PARTICLE is 60 bytes");
static_assert(sizeof(cuco_checkpoint_header) == 168, "checkpoint header is 168 bytes");

namespace {

constexpr int kTimebase = 1 << 28;  // allvars.h:11

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

}  // namespace

extern "C" {

int cuco_write_snapshot(const char *path, const cuco_header *header,
                        const cuco_particle *P)
{
    if (!path || !header || (header->NumPart > 0 && !P) || header->NumPart < 0) {
        set_error("null argument or negative NumPart");
        return CUCO_ERR_INVALID_ARG;
    }
    File fp;
    fp.f = fopen(path, "wb");
    if (!fp.f) {
        set_error("cannot open snapshot file %s for writing", path);
        return CUCO_ERR_INVALID_ARG;
    }
    const size_t n = (size_t)header->NumPart;
    if (fwrite(header, sizeof(cuco_header), 1, fp.f) != 1 ||
        (n > 0 && fwrite(P, sizeof(cuco_particle), n, fp.f) != n)) {
        set_error("short write to %s", path);
        return CUCO_ERR_INVALID_ARG;
    }
    if (fclose(fp.f) != 0) {
        fp.f = nullptr;
        set_error("closing %s failed", path);
        return CUCO_ERR_INVALID_ARG;
    }
    fp.f = nullptr;
    return CUCO_OK;
}

int cuco_read_snapshot(const char *path, cuco_header *header, cuco_particle *P,
                       int64_t capacity)
{
    if (!path || !header) {
        set_error("null argument");
        return CUCO_ERR_INVALID_ARG;
    }
    File fp;
    fp.f = fopen(path, "rb");
    if (!fp.f) {
        set_error("cannot open snapshot file %s", path);
        return CUCO_ERR_INVALID_ARG;
    }
    cuco_header h;
    if (fread(&h, sizeof(h), 1, fp.f) != 1 || h.NumPart < 0) {
        set_error("%s: no valid HEADER record", path);
        return CUCO_ERR_INVALID_ARG;
    }
    if (fseek(fp.f, 0, SEEK_END) != 0) {
        set_error("%s: cannot seek", path);
        return CUCO_ERR_INVALID_ARG;
    }
    const long size = ftell(fp.f);
    const long want = (long)sizeof(cuco_header) + (long)h.NumPart * (long)sizeof(cuco_particle);
    if (size != want) {
        set_error("%s: %ld bytes, but HEADER.NumPart=%d needs %ld", path, size,
                  h.NumPart, want);
        return CUCO_ERR_INVALID_ARG;
    }
    *header = h;
    if (!P) return CUCO_OK;
    if (capacity < h.NumPart) {
        set_error("%s holds %d particles, buffer only %lld", path, h.NumPart,
                  (long long)capacity);
        return CUCO_ERR_INVALID_ARG;
    }
    if (fseek(fp.f, (long)sizeof(cuco_header), SEEK_SET) != 0 ||
        (h.NumPart > 0 &&
         fread(P, sizeof(cuco_particle), (size_t)h.NumPart, fp.f) != (size_t)h.NumPart)) {
        set_error("%s: short read", path);
        return CUCO_ERR_INVALID_ARG;
    }
    return CUCO_OK;
}

int cuco_read_output_list(const char *path, double *times, int32_t max_times,
                          int32_t *count)
{
    if (!path || !count || max_times < 0 || (max_times > 0 && !times)) {
        set_error("null argument");
        return CUCO_ERR_INVALID_ARG;
    }
    File fp;
    fp.f = fopen(path, "r");
    if (!fp.f) {
        set_error("cannot read output list file %s", path);
        return CUCO_ERR_INVALID_ARG;
    }
    int n = 0;
    float v;
    while (n < max_times && fscanf(fp.f, " %g ", &v) == 1) times[n++] = v;
    *count = n;
    return CUCO_OK;
}

int32_t cuco_find_next_output_time(const double *times, int32_t count,
                                   double time_begin, double time_max,
                                   double timebase_interval, int32_t ti_curr)
{
    int ti_next = -1;
    for (int i = 0; i < count; ++i) {
        const double time = times[i];
        if (time >= time_begin && time <= time_max) {
            const int ti = (int)(log(time / time_begin) / timebase_interval);
            if (ti > ti_curr) {
                if (ti_next == -1) ti_next = ti;
                if (ti_next > ti) ti_next = ti;
            }
        }
    }
    if (ti_next == -1) ti_next = 2 * kTimebase;
    return ti_next;
}

int cuco_init_records(cuco_particle *P, int64_t n, double time_begin)
{
    if ((!P && n > 0) || n < 0 || !(time_begin > 0)) {
        set_error("null particle array, n < 0 or time_begin <= 0");
        return CUCO_ERR_INVALID_ARG;
    }
    // init.c:27-36: the IC holds peculiar velocities; the run integrates
    // a^(3/2) times them.  Vel is float, the factor double (float *= double).
    const double fac = sqrt(time_begin) * time_begin;
    for (int64_t i = 0; i < n; ++i) {
        for (int j = 0; j < 3; ++j) {
            P[i].Vel[j] = (float)(P[i].Vel[j] * fac);
            P[i].GravAccel[j] = 0;
            P[i].GravPM[j] = 0;
        }
        P[i].Ti_endstep = 0;
        P[i].Ti_begstep = 0;
    }
    return CUCO_OK;
}

int64_t cuco_checkpoint_bytes(const cuco_checkpoint_header *h)
{
    auto prec_ok = [](int32_t p) { return p == -1 || p == CUCO_FP64 || p == CUCO_FP32; };
    if (!h || h->nboxes < 1 || h->nboxes > (1 << 20) || h->n < 1 || h->n > (int64_t)UINT32_MAX ||
        (h->timeline != 0 && h->timeline != 1) || !prec_ok(h->pp_precision) ||
        !prec_ok(h->pm_precision))
        return -1;
    auto acc = [](int32_t p) -> int64_t { return p < 0 ? 0 : (p == CUCO_FP64 ? 32 : 16); };
    const int64_t per_particle = 2 * 32 + (h->timeline ? 8 : 0) + acc(h->pp_precision) +
                                 acc(h->pm_precision);
    return (int64_t)sizeof(cuco_checkpoint_header) +
           (int64_t)h->nboxes * (int64_t)(sizeof(cuco_params) + sizeof(int64_t)) +
           h->n * per_particle;
}

int cuco_checkpoint_info(const char *path, cuco_checkpoint_header *header)
{
    if (!path || !header) {
        set_error("null argument");
        return CUCO_ERR_INVALID_ARG;
    }
    File fp;
    fp.f = fopen(path, "rb");
    if (!fp.f) {
        set_error("cannot open checkpoint file %s", path);
        return CUCO_ERR_INVALID_ARG;
    }
    cuco_checkpoint_header h;
    if (fread(&h, sizeof(h), 1, fp.f) != 1 || memcmp(h.magic, "CUCOCKPT", 8) != 0) {
        set_error("%s: not a checkpoint file", path);
        return CUCO_ERR_INVALID_ARG;
    }
    if (h.version != 1 || h.reserved != 0) {
        set_error("%s: checkpoint version %d is not supported", path, h.version);
        return CUCO_ERR_INVALID_ARG;
    }
    const int64_t want = cuco_checkpoint_bytes(&h);
    if (want < 0) {
        set_error("%s: header fields out of range", path);
        return CUCO_ERR_INVALID_ARG;
    }
    if (fseek(fp.f, 0, SEEK_END) != 0) {
        set_error("%s: cannot seek", path);
        return CUCO_ERR_INVALID_ARG;
    }
    const long size = ftell(fp.f);
    if ((int64_t)size != want) {
        set_error("%s: %ld bytes, but its header needs %lld", path, size, (long long)want);
        return CUCO_ERR_INVALID_ARG;
    }
    *header = h;
    return CUCO_OK;
}

}  // extern "C"
