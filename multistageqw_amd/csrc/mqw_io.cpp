// Reference file surface (MultiQW.cpp:264-271,303-304,457): instance records of
// n(n+1)/2 little-endian float64 read from an offset, and results appended as
// float32 to output_<n>_<m>.  Short or missing input is an error here; the
// reference would repeat its last record.
#include <errno.h>
#include <fcntl.h>
#include <stdio.h>
#include <string.h>
#include <sys/file.h>
#include <unistd.h>

#include "mqw_internal.h"

extern "C" {

mqw_status mqw_count_instances(const char* path, int n, long* count)
{
    if (!path || !count || n < 2 || n > MQW_MAX_N) return MQW_EINVAL;
    FILE* f = fopen(path, "rb");
    if (!f) return MQW_EIO;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return MQW_EIO;
    }
    const long size = ftell(f);
    fclose(f);
    if (size < 0) return MQW_EIO;
    *count = size / (long)(8 * (n * (n + 1) / 2));
    return MQW_OK;
}

mqw_status mqw_read_instances(const char* path, int n, long start, long count, double* out)
{
    if (!path || n < 2 || n > MQW_MAX_N || start < 0 || count < 0 || (count > 0 && !out)) return MQW_EINVAL;
    const long P = n * (n + 1) / 2;
    FILE* f = fopen(path, "rb");
    if (!f) return MQW_EIO;
    if (fseek(f, 8 * P * start, SEEK_SET) != 0) {
        fclose(f);
        return MQW_EIO;
    }
    const size_t want = (size_t)(P * count);
    const size_t got = fread(out, sizeof(double), want, f);
    fclose(f);
    return got == want ? MQW_OK : MQW_EIO;
}

// One block per call, whole: the file is opened for appending and held under
// an exclusive advisory lock while the block is written, so blocks from
// concurrent slices of a run (run.sh starts one process per slice) land
// contiguously, in completion order, as the reference's single write does.
mqw_status mqw_append_output(const char* dir, int n, int m, const float* probs, long count)
{
    if (count < 0 || (count > 0 && !probs)) return MQW_EINVAL;
    char path[4096];
    snprintf(path, sizeof path, "%s%soutput_%d_%d", dir && *dir ? dir : "", dir && *dir ? "/" : "", n, m);
    const int fd = open(path, O_WRONLY | O_APPEND | O_CREAT, 0644);
    if (fd < 0) return MQW_EIO;
    if (flock(fd, LOCK_EX) != 0) {
        close(fd);
        return MQW_EIO;
    }
    const char* p = (const char*)probs;
    size_t left = sizeof(float) * (size_t)count;
    bool ok = true;
    while (left > 0) {
        const ssize_t w = write(fd, p, left);
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) {
            ok = false;
            break;
        }
        p += w;
        left -= (size_t)w;
    }
    flock(fd, LOCK_UN);
    if (close(fd) != 0) ok = false;
    return ok ? MQW_OK : MQW_EIO;
}

} // extern "C"
