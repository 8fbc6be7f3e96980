// vf_serve.hip -- the serving loop: resident reconstruction streams and
// continuous batching over the fused frame (vf_warped_integrate_batch).
//
// The reference integrates one volume per process, frame after frame, from
// one worker thread (src/VolumeFusion.cu:223-760; src/ui/MainWindow.cpp:234-245).
// Here many streams share the device:
//   - each open stream owns its tsdf/weight volume, node positions, radii,
//     transforms and depth buffer on the device for as long as it is open;
//   - submit() copies a frame's host depth and transforms into pinned
//     staging memory and queues it;
//   - the dispatcher thread takes, oldest first, up to max_batch streams that
//     have a queued frame and no frame in flight, uploads their inputs on a
//     free lane (a HIP stream with its own workspace) and launches one batched
//     frame there;
//   - the completion thread polls the lanes' events, stamps the frames and
//     frees the lanes and streams for the next batch.
// A stream has at most one frame in flight, so its frames integrate in
// submission order and its device input buffers are never overwritten while
// a kernel reads them; batches on different lanes touch disjoint volumes and
// may run concurrently.
#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "vf_host.hpp"

namespace vf {
namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint64_t voxel_count(const vf_int3& d) {
    return (uint64_t)std::max(d.x, 0) * (uint64_t)std::max(d.y, 0) * (uint64_t)std::max(d.z, 0);
}

// Pinned host memory for one queued frame.
struct Staging {
    void* mem = nullptr;
    uint64_t bytes = 0;
};

struct Pending {
    uint64_t ticket;
    int32_t sid;
    double submit_s;
    Staging* stg;       // owned staging, or NULL for a borrowed frame
    const void* depth;  // where the inputs are read from at upload
    const void* T;
};

struct StreamState {
    bool open = false;
    vf_stream_desc desc;
    float* tsdf = nullptr;
    float* wt = nullptr;
    vf_float3* points = nullptr;
    float* radii = nullptr;
    vf_mat4* T = nullptr;
    float* depth = nullptr;
    uint64_t own = 0;  // its buffers
    std::array<uint64_t, kWsGroups> ws{};  // what it adds to each buffer group of a lane workspace
    std::deque<Pending> queue;
    bool busy = false;  // a frame in flight
};

struct Lane {
    hipStream_t stream = nullptr;
    vf_workspace* ws = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;    // a batch formed on this lane and not yet completed
    bool issued = false;  // ... and handed to the device (set under the lock)
    std::vector<Pending> frames;
    std::vector<StreamState*> states;  // frames[i]'s stream, fixed while busy
    double issue_s = 0.0;
    uint64_t batch = 0;
    int status = VF_OK;
};

void free_stream(StreamState& s) {
    void* ptrs[] = {s.tsdf, s.wt, s.points, s.radii, s.T, s.depth};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    s.tsdf = s.wt = s.radii = s.depth = nullptr;
    s.points = nullptr;
    s.T = nullptr;
    s.open = false;
}

}  // namespace
}  // namespace vf

using namespace vf;

struct vf_server {
    vf_server_config cfg;
    std::mutex mu;
    std::condition_variable cv_dispatch;  // work or a free lane appeared, or stop
    std::condition_variable cv_done;      // frames completed, streams closed, lanes issued
    std::vector<std::unique_ptr<StreamState>> streams;  // states never move
    std::vector<Lane> lanes;
    std::vector<Staging*> staging_free;
    std::vector<Staging*> staging_all;
    std::map<uint64_t, vf_frame_result> done_by_ticket;
    std::deque<vf_frame_result> done_log;
    uint64_t next_ticket = 1;
    uint64_t next_batch = 0;
    uint64_t submitted = 0, completed = 0, batches = 0;
    uint64_t bytes_in_use = 0;            // the admission charge of the open (and opening) streams
    uint64_t own_total = 0;               // their buffers
    std::array<std::multiset<uint64_t>, kWsGroups> ws_open;  // their workspace shares, by buffer group
    int max_batch_seen = 0;
    int inflight = 0;
    bool stop = false;
    bool dispatcher_done = false;
    std::thread dispatcher, completer;
};

namespace vf {
namespace {

constexpr size_t kDoneKeep = 1u << 20;  // completed results kept for vf_server_wait

bool has_ready_stream(const vf_server* sv) {
    for (const auto& s : sv->streams)
        if (s->open && !s->busy && !s->queue.empty()) return true;
    return false;
}

int free_lane(const vf_server* sv) {
    for (size_t i = 0; i < sv->lanes.size(); ++i)
        if (!sv->lanes[i].busy) return (int)i;
    return -1;
}

// Uploads the batch's inputs and launches its frame on the lane's stream.
int issue_batch(vf_server* sv, Lane& lane) {
    const int n = (int)lane.frames.size();
    std::vector<vf_volume> vols(n);
    std::vector<vf_graph> graphs(n);
    std::vector<vf_frame> frames(n);
    std::vector<vf_warp_params> params(n);
    for (int i = 0; i < n; ++i) {
        const Pending& f = lane.frames[i];
        StreamState& s = *lane.states[i];  // busy streams are neither closed nor changed
        const vf_stream_desc& d = s.desc;
        const uint64_t depth_bytes = (uint64_t)d.width * (uint64_t)d.height * sizeof(float);
        const uint64_t t_bytes = (uint64_t)d.nodes * sizeof(vf_mat4);
        if (depth_bytes) VF_HIP(hipMemcpyAsync(s.depth, f.depth, depth_bytes, hipMemcpyDefault, lane.stream));
        if (t_bytes) VF_HIP(hipMemcpyAsync(s.T, f.T, t_bytes, hipMemcpyDefault, lane.stream));
        vf_volume& v = vols[i];
        v.tsdf = s.tsdf;
        v.weights = s.wt;
        v.dim = d.dim;
        v.voxel_size = vf_float3{d.size.x / (float)d.dim.x, d.size.y / (float)d.dim.y, d.size.z / (float)d.dim.z};
        v.size = d.size;
        v.center = d.center;
        graphs[i] = vf_graph{s.points, s.points, s.T, s.radii, (size_t)d.nodes};
        frames[i] = vf_frame{s.depth, d.width, d.height, d.intrinsics};
        params[i] = vf_warp_params{sv->cfg.k, d.delta, d.max_weight, d.voxels_per_node, d.blend};
    }
    int st = vf_warped_integrate_batch(lane.ws, n, vols.data(), graphs.data(), frames.data(), params.data(),
                                       lane.stream);
    if (st) return st;
    VF_HIP(hipEventRecord(lane.done, lane.stream));
    return VF_OK;
}

void dispatcher_main(vf_server* sv) {
    (void)hipSetDevice(sv->cfg.device);
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        sv->cv_dispatch.wait(lk, [&] { return sv->stop || (free_lane(sv) >= 0 && has_ready_stream(sv)); });
        if (!(free_lane(sv) >= 0 && has_ready_stream(sv))) {
            // stop requested: finish what is queued before leaving
            bool queued = false;
            for (const auto& s : sv->streams) queued = queued || (s->open && !s->queue.empty());
            if (!queued) break;
            sv->cv_dispatch.wait(lk, [&] { return free_lane(sv) >= 0 && has_ready_stream(sv); });
        }
        const int li = free_lane(sv);
        Lane& lane = sv->lanes[li];
        // streams with a queued frame and none in flight, oldest frame first
        std::vector<std::pair<double, int32_t>> ready;
        for (size_t i = 0; i < sv->streams.size(); ++i) {
            const StreamState& s = *sv->streams[i];
            if (s.open && !s.busy && !s.queue.empty()) ready.emplace_back(s.queue.front().submit_s, (int32_t)i);
        }
        std::sort(ready.begin(), ready.end());
        const size_t take = std::min<size_t>(ready.size(), (size_t)sv->cfg.max_batch);
        lane.frames.clear();
        lane.states.clear();
        for (size_t i = 0; i < take; ++i) {
            StreamState& s = *sv->streams[ready[i].second];
            lane.frames.push_back(s.queue.front());
            lane.states.push_back(&s);
            s.queue.pop_front();
            s.busy = true;
        }
        lane.busy = true;
        lane.batch = sv->next_batch++;
        sv->max_batch_seen = std::max(sv->max_batch_seen, (int)take);
        lk.unlock();
        const double t_issue = now_s();
        const int status = issue_batch(sv, lane);
        lk.lock();
        lane.issue_s = t_issue;
        lane.status = status;
        lane.issued = true;
        sv->inflight++;
        sv->cv_done.notify_all();
    }
    sv->dispatcher_done = true;
    sv->cv_done.notify_all();
}

void completer_main(vf_server* sv) {
    (void)hipSetDevice(sv->cfg.device);
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        sv->cv_done.wait(lk, [&] { return sv->inflight > 0 || sv->dispatcher_done; });
        if (sv->inflight == 0 && sv->dispatcher_done) break;
        // poll the busy lanes (a lane is busy from batch formation until its
        // completion is recorded; its event is recorded once issued)
        std::vector<int> issued;
        for (size_t i = 0; i < sv->lanes.size(); ++i)
            if (sv->lanes[i].issued) issued.push_back((int)i);
        lk.unlock();
        std::vector<std::pair<int, double>> finished;
        for (int li : issued) {
            Lane& lane = sv->lanes[li];
            if (lane.status != VF_OK) {
                finished.emplace_back(li, now_s());
                continue;
            }
            const hipError_t q = hipEventQuery(lane.done);
            if (q == hipSuccess) {
                finished.emplace_back(li, now_s());
            } else if (q != hipErrorNotReady) {
                lane.status = VF_ERR_HIP;
                finished.emplace_back(li, now_s());
            }
        }
        if (finished.empty()) std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
        for (const auto& fd : finished) {
            Lane& lane = sv->lanes[fd.first];
            for (const Pending& f : lane.frames) {
                vf_frame_result r;
                memset(&r, 0, sizeof(r));
                r.ticket = f.ticket;
                r.stream_id = f.sid;
                r.status = lane.status;
                r.submit_s = f.submit_s;
                r.issue_s = lane.issue_s;
                r.done_s = fd.second;
                r.batch = lane.batch;
                r.batch_frames = (uint32_t)lane.frames.size();
                r.lane = (uint32_t)fd.first;
                sv->done_by_ticket[r.ticket] = r;
                while (sv->done_by_ticket.size() > kDoneKeep) sv->done_by_ticket.erase(sv->done_by_ticket.begin());
                sv->done_log.push_back(r);
                sv->streams[f.sid]->busy = false;
                if (f.stg) sv->staging_free.push_back(f.stg);
                sv->completed++;
            }
            lane.frames.clear();
            lane.states.clear();
            lane.busy = false;
            lane.issued = false;
            lane.issue_s = 0.0;
            sv->batches++;
            sv->inflight--;
        }
        if (!finished.empty()) {
            sv->cv_dispatch.notify_all();
            sv->cv_done.notify_all();
        }
    }
}

Staging* take_staging(vf_server* sv, uint64_t bytes) {
    // caller holds sv->mu
    for (size_t i = 0; i < sv->staging_free.size(); ++i) {
        if (sv->staging_free[i]->bytes >= bytes) {
            Staging* s = sv->staging_free[i];
            sv->staging_free.erase(sv->staging_free.begin() + (long)i);
            return s;
        }
    }
    return nullptr;
}

// A stream's buffers: tsdf and weight, node positions, radii, transforms,
// depth frame.
uint64_t stream_own_bytes(const vf_stream_desc& d) {
    const uint64_t n = voxel_count(d.dim);
    return 8 * n + (uint64_t)d.nodes * (12 + 4 + 64) + (uint64_t)std::max(d.width, 0) * std::max(d.height, 0) * 4;
}

// The lanes' pending-record budget, in records per voxel.  On the first
// frames of the 32-stream workload (129^3..491^3 volumes) 1.5 % to 3.2 % of
// the voxels had a pending update; bricks beyond the budget are recomputed
// by the redo pass, so the figure sets speed and memory, not results.
constexpr float kServeRecordsPerVoxel = 1.0f / 24.0f;

// What a stream adds to each buffer group of a lane workspace whose batch
// holds it (lane workspaces use the default brick shape).
std::array<uint64_t, kWsGroups> stream_ws_groups(const vf_stream_desc& d) {
    std::array<uint64_t, kWsGroups> g{};
    volume_ws_bytes(d.dim, d.nodes, d.voxels_per_node, d.width, d.height, 0, kServeRecordsPerVoxel, g.data());
    return g;
}

uint64_t stream_ws_bytes(const vf_stream_desc& d) {
    uint64_t t = 0;
    for (uint64_t x : stream_ws_groups(d)) t += x;
    return t;
}

uint64_t stream_bytes(const vf_stream_desc& d, int lanes) {
    return stream_own_bytes(d) + (uint64_t)std::max(lanes, 1) * stream_ws_bytes(d);
}

// The server's charge: every open stream's buffers, and `lanes` workspaces.
// A lane workspace grows to exactly what its batches need, and a batch holds
// at most max_batch streams, so each of its buffer groups stays within the
// sum of the max_batch largest shares of that group.  Caller holds sv->mu.
uint64_t server_charge(const vf_server* sv, uint64_t own_total,
                       const std::array<std::multiset<uint64_t>, kWsGroups>& ws) {
    uint64_t top = 0;
    for (const auto& grp : ws) {
        int k = 0;
        for (auto it = grp.rbegin(); it != grp.rend() && k < sv->cfg.max_batch; ++it, ++k) top += *it;
    }
    return own_total + (uint64_t)std::max(sv->cfg.lanes, 1) * top;
}

void ws_insert(std::array<std::multiset<uint64_t>, kWsGroups>& ws, const std::array<uint64_t, kWsGroups>& g) {
    for (int i = 0; i < kWsGroups; ++i) ws[i].insert(g[i]);
}

void ws_erase(std::array<std::multiset<uint64_t>, kWsGroups>& ws, const std::array<uint64_t, kWsGroups>& g) {
    for (int i = 0; i < kWsGroups; ++i) ws[i].erase(ws[i].find(g[i]));
}

}  // namespace
}  // namespace vf

extern "C" {

uint64_t vf_server_stream_bytes(const vf_stream_desc* desc, int lanes) { return desc ? stream_bytes(*desc, lanes) : 0; }

int vf_server_create(vf_server** out, const vf_server_config* cfg) {
    if (!out || !cfg) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_create: NULL argument");
    *out = nullptr;
    int st = check_k(cfg->k);
    if (st) return st;
    if (cfg->max_batch < 1 || cfg->max_batch > 256)
        return fail(VF_ERR_INVALID_ARGUMENT, "max_batch %d: supported 1..256", cfg->max_batch);
    if (cfg->lanes < 1 || cfg->lanes > 8) return fail(VF_ERR_INVALID_ARGUMENT, "lanes %d: supported 1..8", cfg->lanes);
    VF_HIP(hipSetDevice(cfg->device));
    vf_server* sv = new vf_server();
    sv->cfg = *cfg;
    sv->lanes.resize(cfg->lanes);
    for (Lane& l : sv->lanes) {
        hipError_t e = hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&l.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            for (Lane& m : sv->lanes) {
                if (m.done) (void)hipEventDestroy(m.done);
                if (m.stream) (void)hipStreamDestroy(m.stream);
                if (m.ws) vf_workspace_destroy(m.ws);
            }
            delete sv;
            return fail(VF_ERR_HIP, "server lane setup: %s", hipGetErrorString(e));
        }
        st = vf_workspace_create(&l.ws, cfg->max_batch, 0, 0, l.stream);
        if (!st) st = vf_workspace_set_record_budget(l.ws, kServeRecordsPerVoxel);
        if (st) {
            for (Lane& m : sv->lanes) {
                if (m.done) (void)hipEventDestroy(m.done);
                if (m.stream) (void)hipStreamDestroy(m.stream);
                if (m.ws) vf_workspace_destroy(m.ws);
            }
            delete sv;
            return st;
        }
    }
    sv->dispatcher = std::thread(dispatcher_main, sv);
    sv->completer = std::thread(completer_main, sv);
    *out = sv;
    return VF_OK;
}

int vf_server_destroy(vf_server* sv) {
    if (!sv) return VF_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_dispatch.notify_all();
    sv->dispatcher.join();
    sv->cv_done.notify_all();
    sv->completer.join();
    (void)hipSetDevice(sv->cfg.device);
    for (Lane& l : sv->lanes) {
        (void)hipStreamSynchronize(l.stream);
        vf_workspace_destroy(l.ws);
        (void)hipEventDestroy(l.done);
        (void)hipStreamDestroy(l.stream);
    }
    for (auto& s : sv->streams)
        if (s->open) free_stream(*s);
    for (Staging* s : sv->staging_all) {
        (void)hipHostFree(s->mem);
        delete s;
    }
    delete sv;
    return VF_OK;
}

int vf_server_open(vf_server* sv, const vf_stream_desc* d, int wait_ms, int32_t* stream_id) {
    if (!sv || !d || !stream_id) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_open: NULL argument");
    *stream_id = -1;
    uint64_t n;
    int st = check_volume(d->dim, vf_float3{d->size.x / (float)std::max(d->dim.x, 1),
                                            d->size.y / (float)std::max(d->dim.y, 1),
                                            d->size.z / (float)std::max(d->dim.z, 1)},
                          &n);
    if (st) return st;
    if (d->width < 0 || d->height < 0) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_open: negative frame size");
    if (d->nodes > 0 && (!d->points || !d->radii))
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_open: node positions and radii must not be NULL");
    const uint64_t own = stream_own_bytes(*d);
    const std::array<uint64_t, kWsGroups> wsb = stream_ws_groups(*d);
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        uint64_t charge = 0;
        const auto fits = [&] {
            auto with = sv->ws_open;
            ws_insert(with, wsb);
            charge = server_charge(sv, sv->own_total + own, with);
            return charge <= sv->cfg.hbm_budget;
        };
        if (!fits()) {
            if (wait_ms > 0) sv->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), fits);
            if (!fits())
                return fail(VF_ERR_CAPACITY, "stream would bring the charge to %llu B; %llu of %llu B in use",
                            (unsigned long long)charge, (unsigned long long)sv->bytes_in_use,
                            (unsigned long long)sv->cfg.hbm_budget);
        }
        // reserved while allocating
        sv->own_total += own;
        ws_insert(sv->ws_open, wsb);
        sv->bytes_in_use = charge;
    }
    const auto unreserve = [&] {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->own_total -= own;
        ws_erase(sv->ws_open, wsb);
        sv->bytes_in_use = server_charge(sv, sv->own_total, sv->ws_open);
        sv->cv_done.notify_all();
    };
    StreamState s;
    s.open = true;
    s.desc = *d;
    s.desc.points = nullptr;
    s.desc.radii = nullptr;
    s.own = own;
    s.ws = wsb;
    const uint64_t g = d->nodes;
    const uint64_t px = (uint64_t)d->width * (uint64_t)d->height;
    hipError_t e = hipSetDevice(sv->cfg.device);
    if (e == hipSuccess) e = hipMalloc((void**)&s.tsdf, std::max<uint64_t>(n, 1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&s.wt, std::max<uint64_t>(n, 1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&s.points, std::max<uint64_t>(g, 1) * sizeof(vf_float3));
    if (e == hipSuccess) e = hipMalloc((void**)&s.radii, std::max<uint64_t>(g, 1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&s.T, std::max<uint64_t>(g, 1) * sizeof(vf_mat4));
    if (e == hipSuccess) e = hipMalloc((void**)&s.depth, std::max<uint64_t>(px, 1) * sizeof(float));
    hipStream_t hs = nullptr;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&hs, hipStreamNonBlocking);
    if (e == hipSuccess && g) e = hipMemcpyAsync(s.points, d->points, g * sizeof(vf_float3), hipMemcpyHostToDevice, hs);
    if (e == hipSuccess && g) e = hipMemcpyAsync(s.radii, d->radii, g * sizeof(float), hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemsetAsync(s.wt, 0, n * sizeof(float), hs);
    int fst = VF_OK;
    if (e == hipSuccess) fst = vf_fill_f32(s.tsdf, -1.0f, n, hs);  // DeviceTsdf init, include/af/DeviceTsdf.cuh:17-19
    if (e == hipSuccess && !fst) e = hipStreamSynchronize(hs);
    if (hs) (void)hipStreamDestroy(hs);
    if (e != hipSuccess || fst) {
        free_stream(s);
        unreserve();
        if (fst) return fst;
        return fail(e == hipErrorOutOfMemory ? VF_ERR_OUT_OF_MEMORY : VF_ERR_HIP, "vf_server_open: %s",
                    hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    int32_t id = -1;
    for (size_t i = 0; i < sv->streams.size(); ++i)
        if (!sv->streams[i]->open && !sv->streams[i]->busy && sv->streams[i]->queue.empty()) {
            id = (int32_t)i;
            break;
        }
    if (id < 0) {
        sv->streams.push_back(std::make_unique<StreamState>());
        id = (int32_t)sv->streams.size() - 1;
    }
    *sv->streams[id] = std::move(s);
    *stream_id = id;
    return VF_OK;
}

int vf_server_close(vf_server* sv, int32_t sid) {
    if (!sv) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_close: NULL server");
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sid < 0 || (size_t)sid >= sv->streams.size() || !sv->streams[sid]->open)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_close: stream %d is not open", sid);
    sv->cv_done.wait(lk, [&] { return !sv->streams[sid]->busy && sv->streams[sid]->queue.empty(); });
    StreamState& s = *sv->streams[sid];
    (void)hipSetDevice(sv->cfg.device);
    free_stream(s);
    sv->own_total -= s.own;
    ws_erase(sv->ws_open, s.ws);
    sv->bytes_in_use = server_charge(sv, sv->own_total, sv->ws_open);
    s.own = 0;
    s.ws = {};
    sv->cv_done.notify_all();
    return VF_OK;
}

}  // extern "C"

namespace vf {
namespace {

int submit_frame(vf_server* sv, int32_t sid, const float* depth, const vf_mat4* transforms, bool borrow,
                 uint64_t* ticket) {
    if (!sv || !ticket) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_submit: NULL argument");
    const double t0 = now_s();
    uint64_t depth_bytes, t_bytes;
    Staging* stg = nullptr;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sid < 0 || (size_t)sid >= sv->streams.size() || !sv->streams[sid]->open)
            return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_submit: stream %d is not open", sid);
        const vf_stream_desc& d = sv->streams[sid]->desc;
        depth_bytes = (uint64_t)d.width * (uint64_t)d.height * sizeof(float);
        t_bytes = (uint64_t)d.nodes * sizeof(vf_mat4);
        if ((depth_bytes && !depth) || (t_bytes && !transforms))
            return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_submit: NULL depth or transforms");
        if (!borrow) stg = take_staging(sv, depth_bytes + t_bytes);
    }
    const void* dsrc = depth;
    const void* tsrc = transforms;
    if (!borrow) {
        if (!stg) {
            stg = new Staging();
            stg->bytes = std::max<uint64_t>(depth_bytes + t_bytes, 1);
            const hipError_t e = hipHostMalloc(&stg->mem, stg->bytes, hipHostMallocDefault);
            if (e != hipSuccess) {
                delete stg;
                return fail(VF_ERR_OUT_OF_MEMORY, "pinned staging of %llu B: %s",
                            (unsigned long long)(depth_bytes + t_bytes), hipGetErrorString(e));
            }
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->staging_all.push_back(stg);
        }
        if (depth_bytes) memcpy(stg->mem, depth, depth_bytes);
        if (t_bytes) memcpy((char*)stg->mem + depth_bytes, transforms, t_bytes);
        dsrc = stg->mem;
        tsrc = (const char*)stg->mem + depth_bytes;
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (!sv->streams[sid]->open) {
            if (stg) sv->staging_free.push_back(stg);
            return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_submit: stream %d closed meanwhile", sid);
        }
        const uint64_t t = sv->next_ticket++;
        sv->streams[sid]->queue.push_back(Pending{t, sid, t0, stg, dsrc, tsrc});
        sv->submitted++;
        *ticket = t;
    }
    sv->cv_dispatch.notify_one();
    return VF_OK;
}

}  // namespace
}  // namespace vf

extern "C" {

int vf_server_submit(vf_server* sv, int32_t sid, const float* depth, const vf_mat4* transforms, uint64_t* ticket) {
    return submit_frame(sv, sid, depth, transforms, false, ticket);
}

int vf_server_submit_borrowed(vf_server* sv, int32_t sid, const float* depth, const vf_mat4* transforms,
                              uint64_t* ticket) {
    return submit_frame(sv, sid, depth, transforms, true, ticket);
}

int vf_server_wait(vf_server* sv, uint64_t ticket, vf_frame_result* result) {
    if (!sv) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_wait: NULL server");
    std::unique_lock<std::mutex> lk(sv->mu);
    if (ticket == 0 || ticket >= sv->next_ticket) return fail(VF_ERR_INVALID_ARGUMENT, "unknown ticket %llu",
                                                              (unsigned long long)ticket);
    sv->cv_done.wait(lk, [&] { return sv->done_by_ticket.count(ticket) > 0 || sv->completed >= sv->submitted; });
    auto it = sv->done_by_ticket.find(ticket);
    if (it == sv->done_by_ticket.end()) return fail(VF_ERR_INVALID_ARGUMENT, "ticket %llu was already collected",
                                                    (unsigned long long)ticket);
    if (result) *result = it->second;
    const int st = it->second.status;
    sv->done_by_ticket.erase(it);
    if (st) return fail(st, "frame of ticket %llu failed", (unsigned long long)ticket);
    return VF_OK;
}

int vf_server_poll(vf_server* sv, vf_frame_result* results, int capacity, int* count) {
    if (!sv || !count || (capacity > 0 && !results)) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_poll: NULL argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    int c = 0;
    while (c < capacity && !sv->done_log.empty()) {
        results[c++] = sv->done_log.front();
        sv->done_log.pop_front();
    }
    *count = c;
    return VF_OK;
}

int vf_server_drain(vf_server* sv) {
    if (!sv) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_drain: NULL server");
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] { return sv->completed >= sv->submitted; });
    return VF_OK;
}

int vf_server_read(vf_server* sv, int32_t sid, float* tsdf, float* weights, vf_float3* points) {
    if (!sv) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_read: NULL server");
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sid < 0 || (size_t)sid >= sv->streams.size() || !sv->streams[sid]->open)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_read: stream %d is not open", sid);
    sv->cv_done.wait(lk, [&] { return !sv->streams[sid]->busy && sv->streams[sid]->queue.empty(); });
    const StreamState& s = *sv->streams[sid];
    const uint64_t n = voxel_count(s.desc.dim);
    VF_HIP(hipSetDevice(sv->cfg.device));
    if (tsdf) VF_HIP(hipMemcpy(tsdf, s.tsdf, n * sizeof(float), hipMemcpyDeviceToHost));
    if (weights) VF_HIP(hipMemcpy(weights, s.wt, n * sizeof(float), hipMemcpyDeviceToHost));
    if (points && s.desc.nodes)
        VF_HIP(hipMemcpy(points, s.points, (uint64_t)s.desc.nodes * sizeof(vf_float3), hipMemcpyDeviceToHost));
    return VF_OK;
}

int vf_server_save(vf_server* sv, int32_t sid, const char* path, int format) {
    if (!sv || !path) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_save: NULL argument");
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sid < 0 || (size_t)sid >= sv->streams.size() || !sv->streams[sid]->open)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_save: stream %d is not open", sid);
    sv->cv_done.wait(lk, [&] { return !sv->streams[sid]->busy && sv->streams[sid]->queue.empty(); });
    const StreamState& s = *sv->streams[sid];
    vf_tsdf_info info;
    memset(&info, 0, sizeof(info));
    info.dim = s.desc.dim;
    info.size = s.desc.size;
    info.center = s.desc.center;
    info.delta = s.desc.delta;
    info.max_weight = s.desc.max_weight;
    VF_HIP(hipSetDevice(sv->cfg.device));
    return vf_tsdf_save(path, format, &info, s.tsdf, s.wt);
}

int vf_server_restore(vf_server* sv, int32_t sid, const char* path) {
    if (!sv || !path) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_restore: NULL argument");
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sid < 0 || (size_t)sid >= sv->streams.size() || !sv->streams[sid]->open)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_restore: stream %d is not open", sid);
    sv->cv_done.wait(lk, [&] { return !sv->streams[sid]->busy && sv->streams[sid]->queue.empty(); });
    const StreamState& s = *sv->streams[sid];
    vf_tsdf_info info;
    int st = vf_tsdf_load_host(path, &info, nullptr, nullptr, 0);
    if (st) return st;
    if (info.dim.x != s.desc.dim.x || info.dim.y != s.desc.dim.y || info.dim.z != s.desc.dim.z)
        return fail(VF_ERR_INVALID_ARGUMENT, "%s holds a %dx%dx%d volume; stream %d is %dx%dx%d", path, info.dim.x,
                    info.dim.y, info.dim.z, sid, s.desc.dim.x, s.desc.dim.y, s.desc.dim.z);
    VF_HIP(hipSetDevice(sv->cfg.device));
    return vf_tsdf_load(path, &info, s.tsdf, s.wt, voxel_count(s.desc.dim));
}

int vf_server_get_stats(vf_server* sv, vf_server_stats* out) {
    if (!sv || !out) return fail(VF_ERR_INVALID_ARGUMENT, "vf_server_get_stats: NULL argument");
    memset(out, 0, sizeof(*out));
    std::vector<Lane*> lanes;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        out->frames_submitted = sv->submitted;
        out->frames_done = sv->completed;
        out->batches = sv->batches;
        out->bytes_in_use = sv->bytes_in_use;
        out->hbm_budget = sv->cfg.hbm_budget;
        out->max_batch_seen = sv->max_batch_seen;
        for (const auto& s : sv->streams) out->streams_open += s->open ? 1 : 0;
        for (Lane& l : sv->lanes) lanes.push_back(&l);
    }
    VF_HIP(hipSetDevice(sv->cfg.device));
    for (Lane* l : lanes) {
        vf_frame_stats t;
        const int st = vf_workspace_totals(l->ws, &t, 0, l->stream);
        if (st) return st;
        out->totals.supported += t.supported;
        out->totals.in_bounds += t.in_bounds;
        out->totals.pending += t.pending;
        out->totals.collided += t.collided;
        out->totals.updated += t.updated;
        out->totals.active_bricks += t.active_bricks;
        out->totals.overflow_bricks += t.overflow_bricks;
        out->totals.redo_bricks += t.redo_bricks;
        out->totals.screened_bricks += t.screened_bricks;
    }
    return VF_OK;
}

}  // extern "C"
