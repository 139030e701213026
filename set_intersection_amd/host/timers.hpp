// Named phase timers with the reference's per-label tables.
//
// HostTimer follows include/host_timer.hpp:11-66 (std::chrono intervals,
// microsecond resolution, reported in ms).  DeviceTimer follows
// include/device_timer.cuh:12-77 but records its events through the C ABI
// (si_event_record / si_event_elapsed) on the context's stream, so the
// executables never call the HIP runtime directly.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "cli.hpp"
#include "set_intersection.h"

namespace si {

inline void printTimings(const char* title, const std::vector<std::string>& names,
                         const std::vector<double>& sums, double total) {
    std::printf("┌%s┐\n", center(title, 51, "─").c_str());
    for (size_t i = 0; i < names.size(); ++i)
        std::printf("│%s|%s│\n", center(names[i], 25).c_str(), center(toText(sums[i]), 25).c_str());
    std::printf("└%s┘\n", center("", 51, "─").c_str());
    std::printf("│%s|%s│\n", center("Total", 25).c_str(), center(toText(total), 25).c_str());
    std::printf("└%s┘\n", center("", 51, "─").c_str());
}

struct Interval {
    using clock = std::chrono::high_resolution_clock;
    clock::time_point begin, end;
    std::string name;
    explicit Interval(std::string n) : name(std::move(n)) {}
};

struct HostTimer {
    std::vector<std::unique_ptr<Interval>> intervals;
    Interval* add(const std::string& name) {
        intervals.emplace_back(new Interval(name));
        intervals.back()->begin = Interval::clock::now();
        return intervals.back().get();
    }
    static void finish(Interval* iv) { iv->end = Interval::clock::now(); }
    static double ms(const Interval& iv) {
        return std::chrono::duration_cast<std::chrono::microseconds>(iv.end - iv.begin).count() / 1000.0;
    }
    double sum(const std::string& name) const {
        double s = 0.0;
        for (auto& iv : intervals)
            if (iv->name == name) s += ms(*iv);
        return s;
    }
    double total() const {
        double s = 0.0;
        for (auto& iv : intervals) s += ms(*iv);
        return s;
    }
    void print() const {
        std::vector<std::string> names;
        for (auto& iv : intervals)
            if (std::find(names.begin(), names.end(), iv->name) == names.end()) names.push_back(iv->name);
        std::vector<double> sums;
        for (auto& n : names) sums.push_back(sum(n));
        printTimings("Host Timings (in ms)", names, sums, total());
    }
};

struct EventPair {
    std::string name;
    int32_t start = -1, end = -1;
    si_context* ctx;
    EventPair(std::string n, si_context* c) : name(std::move(n)), ctx(c) {}
};

struct DeviceTimer {
    si_context* ctx;
    std::vector<std::unique_ptr<EventPair>> pairs;
    explicit DeviceTimer(si_context* c) : ctx(c) {}
    EventPair* add(const std::string& name) {
        pairs.emplace_back(new EventPair(name, ctx));
        si_event_record(ctx, &pairs.back()->start);
        return pairs.back().get();
    }
    static void finish(EventPair* p) { si_event_record(p->ctx, &p->end); }
    // a pair of events the library recorded itself (si_buffer_download_async)
    EventPair* addRecorded(const std::string& name, int32_t start, int32_t end) {
        pairs.emplace_back(new EventPair(name, ctx));
        pairs.back()->start = start;
        pairs.back()->end = end;
        return pairs.back().get();
    }
    // "Transfer to device" and "Compute set offsets" from the events that
    // si_dataset_upload recorded around its copies and its offsets scan
    void addUpload(si_dataset* ds) {
        int32_t ev[3] = {-1, -1, -1};
        si_dataset_upload_events(ds, &ev[0], &ev[1], &ev[2]);
        addRecorded("Transfer to device", ev[0], ev[1]);
        addRecorded("Compute set offsets", ev[1], ev[2]);
    }
    double ms(const EventPair& p) const {
        float v = 0.0f;
        if (p.start < 0 || p.end < 0 || si_event_elapsed(ctx, p.start, p.end, &v) != SI_OK) return 0.0;
        return v;
    }
    double sum(const std::string& name) const {
        double s = 0.0;
        for (auto& p : pairs)
            if (p->name == name) s += ms(*p);
        return s;
    }
    double total() const {
        double s = 0.0;
        for (auto& p : pairs) s += ms(*p);
        return s;
    }
    void print() const {
        std::vector<std::string> names;
        for (auto& p : pairs)
            if (std::find(names.begin(), names.end(), p->name) == names.end()) names.push_back(p->name);
        std::vector<double> sums;
        for (auto& n : names) sums.push_back(sum(n));
        printTimings("Device Timings (in ms)", names, sums, total());
    }
};

}  // namespace si
