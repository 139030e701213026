// Multi-run pipeline of the executables.
//
// The reference's drivers run each tile pair's kernels, copy the run's counts
// into a host vector that holds every run with a blocking cudaMemcpy
// (ip.cu:140-177, matrix_multiplication.cu:138-180), and write the whole
// multi-run file at the end (io.hpp:104-136).  Here run i's counts travel to
// page-locked host memory on the library's copy stream while run i+1 computes,
// and run i's lines are written while the GPU works on run i+1: two device
// count buffers and two host blocks in place of runs x partition^2 counts.
// The file is byte-identical to writeResult's multi-run layout.
//
//   run i:     [kernels i] [copy i]
//   run i+1:               [kernels i+1] [copy i+1]
//   host:                  [wait copy i, write lines of run i] ...
//
// Buffer reuse: run i+2 writes the device block run i used; its kernels are
// enqueued only after the host has waited for copy i (done while writing run
// i's lines), and its copy lands in host block i % 2 after those lines were
// written.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "format.hpp"
#include "set_intersection.h"
#include "tiles.hpp"
#include "timers.hpp"

namespace si {

// launch(iter, run, dst) enqueues run `iter`'s kernels writing its counts to
// dst (a block of partition^2 entries of T) and returns the number of entries
// the run fills (its prefix of the block), or a negative value after a library
// error.  Returns SI_OK or the failing call's status; throws FormatError when
// the output file cannot be written.  hostTimer (optional) times the host
// blocks' allocation as "Allocate memory".
template <typename T, bool mm = false, typename Launch>
int runPipelined(si_context* ctx, const std::vector<tile_pair>& runs, uint32_t partition, si_buffer* const dev[2],
                 DeviceTimer& timer, HostTimer* hostTimer, const std::string& output, Launch&& launch) {
    const uint64_t block = uint64_t(partition) * partition;
    void* host[2] = {nullptr, nullptr};
    struct HostBlocks {
        si_context* ctx;
        void** p;
        ~HostBlocks() {
            si_host_free(ctx, p[0]);
            si_host_free(ctx, p[1]);
        }
    } guard{ctx, host};
    Interval* alloc = hostTimer ? hostTimer->add("Allocate memory") : nullptr;
    for (auto& h : host) {
        int st = si_host_alloc(ctx, block * sizeof(T), &h);
        if (st != SI_OK) return st;
    }
    if (alloc) HostTimer::finish(alloc);
    std::unique_ptr<LineWriter> w;
    if (!output.empty()) w.reset(new LineWriter(output));

    int32_t done[2] = {-1, -1};
    auto drain = [&](size_t iter) -> int {
        const int s = int(iter & 1);
        int st = si_event_wait(ctx, done[s]);
        if (st != SI_OK) return st;
        if (w) writeRun<T, mm>(*w, runs[iter], partition, static_cast<const T*>(host[s]));
        return SI_OK;
    };
    for (size_t iter = 0; iter < runs.size(); ++iter) {
        const int s = int(iter & 1);
        const int64_t used = launch(iter, runs[iter], dev[s]);
        if (used < 0) return SI_ERR_HIP;
        if (uint64_t(used) > block) return SI_ERR_RANGE;
        int32_t start = -1, end = -1;
        int st = si_buffer_download_async(dev[s], 0, host[s], uint64_t(used) * sizeof(T), &start, &end);
        if (st != SI_OK) return st;
        timer.addRecorded("Transfer result", start, end);
        done[s] = end;
        if (iter > 0 && (st = drain(iter - 1)) != SI_OK) return st;
    }
    if (!runs.empty()) {
        int st = drain(runs.size() - 1);
        if (st != SI_OK) return st;
    }
    if (w) w->close();
    return SI_OK;
}

}  // namespace si
