// host_animation.cpp -- FLWA animation files written from device memory, and read back.
//
// The file format is the reference's (src/animation/format.rs:60-233): a 52-byte little-endian
// header, the frames, and a table of (offset, size) entries at the end.  A frame is the state's
// channels back to back as f32, which is exactly the device layout of a context's state
// ([C][D*H*W]), so recording a frame is one device-to-host copy.
//
// Recorder (AnimationRecorder, recorder.rs:47-205).  A frame taken from a context or a
// population member is copied into one of three pinned staging buffers on the owner's stream,
// behind the steps already queued there, and an event marks its completion; a background
// thread waits for the event, compresses the frame if asked (LZ4), writes it and notes its
// index entry.  The caller returns as soon as the copy is queued, so the device goes on
// stepping while earlier frames reach the disk; a fourth frame in flight waits for a free
// buffer.
//
// Player (AnimationPlayer, player.rs:26-200): header, index table, frame reads; and a resume
// path that makes a frame the device-resident state of a context.
//
// LZ4 is the block format with a 4-byte little-endian uncompressed-size prefix (what the
// reference's optional lz4 feature writes through lz4_flex::compress_prepend_size): a greedy
// single-probe hash encoder and a bounds-checked decoder, written from the published block
// format description.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/flowlenia_hip.h"
#include "host_internal.hpp"

namespace {

thread_local std::string g_anim_error;

hp_status afail(hp_status st, const std::string& m) {
    g_anim_error = m;
    return st;
}

constexpr size_t kHeaderSize = 52;  // AnimationHeader::SIZE (format.rs:80-83)
constexpr size_t kIndexSize = 16;   // FrameIndex::SIZE (format.rs:197)
constexpr uint16_t kVersion = 1;

struct Header {
    uint32_t w = 0, h = 0, d = 1, c = 1;
    uint64_t frames = 0;
    float dt = 0.0f;
    uint16_t flags = 0;  // compression in bits 0-3, delta encoding in bit 4 (format.rs:40-56)
    size_t frame_bytes() const { return (size_t)w * h * d * c * 4; }
    int compression() const { return flags & 0x0F; }
};

void put_le(uint8_t* p, uint64_t v, int n) {
    for (int i = 0; i < n; i++) p[i] = (uint8_t)(v >> (8 * i));
}
uint64_t get_le(const uint8_t* p, int n) {
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

void encode_header(const Header& h, uint8_t* b) {  // AnimationHeader::write_to (format.rs:101-115)
    memset(b, 0, kHeaderSize);
    memcpy(b, "FLWA", 4);
    put_le(b + 4, kVersion, 2);
    put_le(b + 6, h.flags, 2);
    put_le(b + 8, h.w, 4);
    put_le(b + 12, h.h, 4);
    put_le(b + 16, h.d, 4);
    put_le(b + 20, h.c, 4);
    put_le(b + 24, h.frames, 8);
    uint32_t dtb;
    memcpy(&dtb, &h.dt, 4);
    put_le(b + 32, dtb, 4);
}

hp_status decode_header(const uint8_t* b, Header* h) {  // AnimationHeader::read_from (format.rs:118-176)
    if (memcmp(b, "FLWA", 4) != 0) return afail(HP_ERR_IO, "Invalid FLWA magic bytes");
    const uint16_t v = (uint16_t)get_le(b + 4, 2);
    if (v != kVersion) return afail(HP_ERR_IO, "Unsupported FLWA version: " + std::to_string(v));
    h->flags = (uint16_t)get_le(b + 6, 2);
    h->w = (uint32_t)get_le(b + 8, 4);
    h->h = (uint32_t)get_le(b + 12, 4);
    h->d = (uint32_t)get_le(b + 16, 4);
    h->c = (uint32_t)get_le(b + 20, 4);
    h->frames = get_le(b + 24, 8);
    const uint32_t dtb = (uint32_t)get_le(b + 32, 4);
    memcpy(&h->dt, &dtb, 4);
    return HP_OK;
}

// ---- LZ4 block format, size-prefixed ------------------------------------------------------
void lz4_len_ext(std::vector<uint8_t>& o, size_t n) {  // length beyond 15: 255-bytes then the rest
    while (n >= 255) {
        o.push_back(255);
        n -= 255;
    }
    o.push_back((uint8_t)n);
}

std::vector<uint8_t> lz4_compress(const uint8_t* src, size_t n) {
    std::vector<uint8_t> o;
    o.reserve(n + n / 255 + 16);
    o.resize(4);
    put_le(o.data(), (uint64_t)n, 4);
    auto literals = [&](size_t from, size_t to, size_t mlen, size_t off) {
        const size_t lit = to - from;
        const size_t ml = mlen ? mlen - 4 : 0;
        o.push_back((uint8_t)(((lit >= 15 ? 15 : lit) << 4) | (mlen ? (ml >= 15 ? 15 : ml) : 0)));
        if (lit >= 15) lz4_len_ext(o, lit - 15);
        o.insert(o.end(), src + from, src + to);
        if (!mlen) return;
        o.push_back((uint8_t)(off & 0xFF));
        o.push_back((uint8_t)(off >> 8));
        if (ml >= 15) lz4_len_ext(o, ml - 15);
    };
    size_t anchor = 0;
    if (n >= 13) {
        // the last match starts at least 12 bytes before the end and ends at least 5 before it
        const size_t limit = n - 12, mlimit = n - 5;
        std::vector<uint32_t> table(1u << 16, 0);  // position + 1 of the last 4-byte sequence
        auto rd32 = [&](size_t i) {
            uint32_t v;
            memcpy(&v, src + i, 4);
            return v;
        };
        size_t ip = 0;
        while (ip < limit) {
            const uint32_t v = rd32(ip);
            const uint32_t hsh = (v * 2654435761u) >> 16;
            const size_t ref = table[hsh];
            table[hsh] = (uint32_t)(ip + 1);
            if (ref && ip - (ref - 1) <= 65535 && rd32(ref - 1) == v) {
                const size_t r = ref - 1;
                size_t ml = 4;
                while (ip + ml < mlimit && src[r + ml] == src[ip + ml]) ml++;
                literals(anchor, ip, ml, ip - r);
                ip += ml;
                anchor = ip;
            } else {
                ip++;
            }
        }
    }
    literals(anchor, n, 0, 0);  // the last sequence holds literals only
    return o;
}

hp_status lz4_decompress(const uint8_t* src, size_t n, std::vector<uint8_t>* out) {
    if (n < 4) return afail(HP_ERR_IO, "LZ4 frame shorter than its size prefix");
    const size_t size = (size_t)get_le(src, 4);
    out->assign(size, 0);
    uint8_t* d = out->data();
    size_t ip = 4, op = 0;
    auto bad = [] { return afail(HP_ERR_IO, "corrupt LZ4 block"); };
    while (ip < n) {
        const uint8_t tok = src[ip++];
        size_t lit = tok >> 4;
        if (lit == 15) {
            uint8_t b;
            do {
                if (ip >= n) return bad();
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (ip + lit > n || op + lit > size) return bad();
        memcpy(d + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip >= n) break;  // the last sequence ends after its literals
        if (ip + 2 > n) return bad();
        const size_t off = (size_t)src[ip] | ((size_t)src[ip + 1] << 8);
        ip += 2;
        if (off == 0 || off > op) return bad();
        size_t ml = (tok & 15) + 4;
        if ((tok & 15) == 15) {
            uint8_t b;
            do {
                if (ip >= n) return bad();
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        if (op + ml > size) return bad();
        for (size_t k = 0; k < ml; k++, op++) d[op] = d[op - off];  // overlapping copies repeat
    }
    if (op != size) return bad();
    return HP_OK;
}

}  // namespace

#define STATUS_OK_OR_RETURN(x)          \
    do {                                \
        const hp_status s_ = (x);       \
        if (s_ != HP_OK) return s_;     \
    } while (0)

struct hp_recorder {
    static constexpr int kSlots = 3;
    FILE* f = nullptr;
    Header hdr;
    hp_recorder_config cfg{};
    uint64_t frames = 0;     // frames accepted (frame_skip / max_frames decisions)
    uint32_t counter = 0;    // calls since the last recorded frame
    std::vector<std::pair<uint64_t, uint64_t>> index;  // written by the writer thread
    uint64_t offset = kHeaderSize;
    // staging: pinned once a device frame is recorded, plain host memory before that
    uint8_t* slot[kSlots] = {};
    bool pinned = false;
    int device = -1;
    hipEvent_t ev[kSlots] = {};
    bool busy[kSlots] = {};
    int next = 0;
    // writer thread
    struct Job {
        int s;
        bool on_device;
    };
    std::thread writer;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> q;
    bool stop = false;
    hp_status wst = HP_OK;
    std::string werr;

    void run() {
        bool device_set = false;
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !q.empty(); });
                if (q.empty()) return;
                j = q.front();
                q.pop_front();
            }
            hp_status st = HP_OK;
            std::string msg;
            if (j.on_device) {
                if (!device_set) {
                    (void)hipSetDevice(device);
                    device_set = true;
                }
                const hipError_t e = hipEventSynchronize(ev[j.s]);
                if (e != hipSuccess) {
                    st = HP_ERR_HIP;
                    msg = std::string("frame copy: ") + hipGetErrorString(e);
                }
            }
            if (st == HP_OK && wst == HP_OK) {
                const size_t nb = hdr.frame_bytes();
                std::vector<uint8_t> packed;
                const uint8_t* data = slot[j.s];
                size_t size = nb;
                if (hdr.compression() == HP_COMPRESSION_LZ4) {
                    packed = lz4_compress(slot[j.s], nb);
                    data = packed.data();
                    size = packed.size();
                }
                if (fwrite(data, 1, size, f) != size) {
                    st = HP_ERR_IO;
                    msg = "short write to the animation file";
                } else {
                    index.push_back({offset, (uint64_t)size});
                    offset += size;
                }
            }
            std::lock_guard<std::mutex> lk(mu);
            if (st != HP_OK && wst == HP_OK) {
                wst = st;
                werr = msg;
            }
            busy[j.s] = false;
            cv.notify_all();
        }
    }

    // the next staging buffer, once the writer has released it
    int take_slot() {
        std::unique_lock<std::mutex> lk(mu);
        const int s = next;
        cv.wait(lk, [&] { return !busy[s]; });
        busy[s] = true;
        next = (next + 1) % kSlots;
        return s;
    }

    void submit(int s, bool on_device) {
        std::lock_guard<std::mutex> lk(mu);
        q.push_back({s, on_device});
        cv.notify_all();
    }

    // frame_skip and max_frames (recorder.rs:119-131): false if this call records nothing
    bool accept() {
        counter++;
        if (counter < std::max<uint32_t>(1, cfg.frame_skip)) return false;
        counter = 0;
        if (cfg.max_frames > 0 && frames >= cfg.max_frames) return false;
        return true;
    }

    hp_status writer_status() {
        std::lock_guard<std::mutex> lk(mu);
        if (wst != HP_OK) return afail(wst, werr);
        return HP_OK;
    }

    void wait_idle() {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] {
            if (!q.empty()) return false;
            for (bool b : busy)
                if (b) return false;
            return true;
        });
    }

    void free_slots() {
        for (int i = 0; i < kSlots; i++) {
            if (slot[i]) {
                if (pinned) (void)hipHostFree(slot[i]);
                else free(slot[i]);
            }
            slot[i] = nullptr;
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            ev[i] = nullptr;
        }
    }

    // pinned staging and events on `dev` (replacing plain host buffers once they are idle)
    hp_status ensure_pinned(int dev) {
        if (pinned) {
            if (dev != device) return afail(HP_ERR_SHAPE, "frames of one recorder must come from one device");
            return HP_OK;
        }
        wait_idle();
        free_slots();
        device = dev;
        if (hipSetDevice(dev) != hipSuccess) return afail(HP_ERR_HIP, "hipSetDevice failed");
        for (int i = 0; i < kSlots; i++) {
            if (hipHostMalloc((void**)&slot[i], std::max<size_t>(1, hdr.frame_bytes()), hipHostMallocDefault) !=
                    hipSuccess ||
                hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
                pinned = true;  // whatever was allocated is freed as pinned memory
                return afail(HP_ERR_HIP, "pinned staging allocation failed");
            }
        }
        pinned = true;
        return HP_OK;
    }

    hp_status ensure_host() {
        if (slot[0]) return HP_OK;
        for (int i = 0; i < kSlots; i++) {
            slot[i] = (uint8_t*)malloc(std::max<size_t>(1, hdr.frame_bytes()));
            if (!slot[i]) return afail(HP_ERR_IO, "out of host memory for frame staging");
        }
        return HP_OK;
    }

    hp_status record_device(const hp::DeviceState& ds, int* recorded) {
        if (recorded) *recorded = 0;
        STATUS_OK_OR_RETURN(writer_status());
        if (ds.W != hdr.w || ds.H != hdr.h || ds.D != hdr.d || ds.C != hdr.c)
            return afail(HP_ERR_SHAPE, "state shape differs from the recorder's header");
        if (!accept()) return HP_OK;
        STATUS_OK_OR_RETURN(ensure_pinned(ds.device));
        const int s = take_slot();
        (void)hipSetDevice(ds.device);
        hipError_t e = hipMemcpyAsync(slot[s], ds.data, hdr.frame_bytes(), hipMemcpyDeviceToHost, ds.stream);
        if (e == hipSuccess) e = hipEventRecord(ev[s], ds.stream);
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            busy[s] = false;
            return afail(HP_ERR_HIP, std::string("frame copy: ") + hipGetErrorString(e));
        }
        submit(s, true);
        frames++;
        if (recorded) *recorded = 1;
        return HP_OK;
    }

    hp_status record_host(const float* const* ch, int* recorded) {
        if (recorded) *recorded = 0;
        STATUS_OK_OR_RETURN(writer_status());
        for (uint32_t c = 0; c < hdr.c; c++)
            if (!ch[c]) return afail(HP_ERR_NULL, "channel pointer is NULL");
        if (!accept()) return HP_OK;
        if (!pinned) STATUS_OK_OR_RETURN(ensure_host());
        const int s = take_slot();
        const size_t nch = (size_t)hdr.w * hdr.h * hdr.d * 4;
        for (uint32_t c = 0; c < hdr.c; c++) memcpy(slot[s] + c * nch, ch[c], nch);
        submit(s, false);
        frames++;
        if (recorded) *recorded = 1;
        return HP_OK;
    }

};

struct hp_player {
    FILE* f = nullptr;
    Header hdr;
    std::vector<std::pair<uint64_t, uint64_t>> index;
    std::vector<uint8_t> buf, raw;

    // AnimationPlayer::read_frame_into (player.rs:130-180) into one contiguous frame
    hp_status read(uint64_t i, std::vector<uint8_t>** frame) {
        if (i >= hdr.frames)
            return afail(HP_ERR_SHAPE, "Frame index " + std::to_string(i) + " out of range (max " +
                                           std::to_string(hdr.frames ? hdr.frames - 1 : 0) + ")");
        const auto& e = index[(size_t)i];
        buf.resize((size_t)e.second);
        if (fseeko(f, (off_t)e.first, SEEK_SET) != 0 || fread(buf.data(), 1, buf.size(), f) != buf.size())
            return afail(HP_ERR_IO, "short read of frame " + std::to_string(i));
        std::vector<uint8_t>* out = &buf;
        if (hdr.compression() == HP_COMPRESSION_LZ4) {
            hp_status st = lz4_decompress(buf.data(), buf.size(), &raw);
            if (st != HP_OK) return st;
            out = &raw;
        }
        if (out->size() != hdr.frame_bytes())
            return afail(HP_ERR_IO, "Frame size mismatch: " + std::to_string(out->size()) + " bytes vs " +
                                        std::to_string(hdr.frame_bytes() / 4) + " floats");
        *frame = out;
        return HP_OK;
    }
};

extern "C" {

const char* hp_animation_last_error(void) { return g_anim_error.c_str(); }

hp_status hp_recorder_open(const char* path, const hp_config* cfg, const hp_recorder_config* rc, hp_recorder** out) {
    if (!out) return afail(HP_ERR_NULL, "out is NULL");
    *out = nullptr;
    if (!path || !cfg) return afail(HP_ERR_NULL, "path or config is NULL");
    const hp_recorder_config c = rc ? *rc : hp_recorder_config{HP_COMPRESSION_NONE, 1, 0};
    if (c.compression != HP_COMPRESSION_NONE && c.compression != HP_COMPRESSION_LZ4)
        return afail(HP_ERR_SHAPE, "unknown compression " + std::to_string(c.compression));
    if (cfg->width == 0 || cfg->height == 0 || cfg->depth == 0 || cfg->channels == 0)
        return afail(HP_ERR_INVALID_DIMENSIONS, "grid dimensions must be non-zero");
    try {
        hp_recorder* r = new hp_recorder();
        r->cfg = c;
        r->hdr.w = cfg->width;
        r->hdr.h = cfg->height;
        r->hdr.d = cfg->depth;
        r->hdr.c = cfg->channels;
        r->hdr.dt = cfg->dt;
        r->hdr.flags = (uint16_t)(c.compression & 0x0F);
        r->f = fopen(path, "wb");
        if (!r->f) {
            delete r;
            return afail(HP_ERR_IO, std::string("cannot create ") + path);
        }
        uint8_t h[kHeaderSize];
        encode_header(r->hdr, h);  // placeholder; finalize rewrites it with the frame count
        if (fwrite(h, 1, kHeaderSize, r->f) != kHeaderSize) {
            fclose(r->f);
            delete r;
            return afail(HP_ERR_IO, "cannot write the header");
        }
        r->writer = std::thread([r] { r->run(); });
        *out = r;
        return HP_OK;
    } catch (const std::exception& e) {
        return afail(HP_ERR_IO, e.what());
    }
}

hp_status hp_recorder_record_ctx(hp_recorder* r, hp_ctx* ctx, int* recorded) {
    if (!r || !ctx) return afail(HP_ERR_NULL, "recorder or context is NULL");
    hp::DeviceState ds;
    hp_status st = hp::ctx_device_state(ctx, &ds);
    if (st != HP_OK) return afail(st, "context state");
    return r->record_device(ds, recorded);
}

hp_status hp_recorder_record_member(hp_recorder* r, hp_batch* b, uint32_t member, int* recorded) {
    if (!r || !b) return afail(HP_ERR_NULL, "recorder or population is NULL");
    hp::DeviceState ds;
    hp_status st = hp::batch_device_state(b, member, &ds);
    if (st != HP_OK) return afail(st, "member index out of range");
    return r->record_device(ds, recorded);
}

hp_status hp_recorder_record_host(hp_recorder* r, const float* const* channels, int* recorded) {
    if (!r || !channels) return afail(HP_ERR_NULL, "recorder or channels is NULL");
    return r->record_host(channels, recorded);
}

hp_status hp_recorder_finalize(hp_recorder* r, hp_animation_stats* stats) {
    if (!r) return afail(HP_ERR_NULL, "recorder is NULL");
    {
        std::lock_guard<std::mutex> lk(r->mu);
        r->stop = true;
        r->cv.notify_all();
    }
    if (r->writer.joinable()) r->writer.join();
    hp_status st = r->wst;
    std::string msg = r->werr;
    if (st == HP_OK) {  // index table at the end, then the header with the frame count
        std::vector<uint8_t> tab(r->index.size() * kIndexSize);
        for (size_t i = 0; i < r->index.size(); i++) {
            put_le(tab.data() + i * kIndexSize, r->index[i].first, 8);
            put_le(tab.data() + i * kIndexSize + 8, r->index[i].second, 8);
        }
        r->hdr.frames = r->index.size();
        uint8_t h[kHeaderSize];
        encode_header(r->hdr, h);
        if (fwrite(tab.data(), 1, tab.size(), r->f) != tab.size() || fseeko(r->f, 0, SEEK_SET) != 0 ||
            fwrite(h, 1, kHeaderSize, r->f) != kHeaderSize) {
            st = HP_ERR_IO;
            msg = "cannot write the index table or the header";
        }
    }
    if (fclose(r->f) != 0 && st == HP_OK) {
        st = HP_ERR_IO;
        msg = "closing the animation file failed";
    }
    if (st == HP_OK && stats) {  // AnimationStats (recorder.rs:188-204)
        const uint64_t n = r->index.size();
        stats->frame_count = n;
        stats->total_bytes = r->offset + n * kIndexSize;
        stats->average_frame_size = n ? (r->offset - kHeaderSize) / n : 0;
        stats->compression = r->hdr.compression();
    }
    r->free_slots();
    delete r;
    return st == HP_OK ? HP_OK : afail(st, msg);
}

hp_status hp_player_open(const char* path, hp_player** out) {
    if (!out) return afail(HP_ERR_NULL, "out is NULL");
    *out = nullptr;
    if (!path) return afail(HP_ERR_NULL, "path is NULL");
    FILE* f = fopen(path, "rb");
    if (!f) return afail(HP_ERR_IO, std::string("cannot open ") + path);
    hp_player* p = new hp_player();
    p->f = f;
    uint8_t h[kHeaderSize];
    hp_status st = HP_OK;
    if (fread(h, 1, kHeaderSize, f) != kHeaderSize) st = afail(HP_ERR_IO, "file shorter than an FLWA header");
    if (st == HP_OK) st = decode_header(h, &p->hdr);
    if (st == HP_OK) {  // the index table closes the file (player.rs:43-55)
        if (fseeko(f, 0, SEEK_END) != 0) st = afail(HP_ERR_IO, "cannot seek");
        const off_t len = ftello(f);
        const uint64_t tab = p->hdr.frames * kIndexSize;
        if (st == HP_OK && (p->hdr.frames > ((uint64_t)len) / kIndexSize || (uint64_t)len < kHeaderSize + tab))
            st = afail(HP_ERR_IO, "file too short for its frame index");
        if (st == HP_OK) {
            std::vector<uint8_t> t((size_t)tab);
            if (fseeko(f, (off_t)((uint64_t)len - tab), SEEK_SET) != 0 || fread(t.data(), 1, t.size(), f) != t.size())
                st = afail(HP_ERR_IO, "short read of the frame index");
            for (uint64_t i = 0; st == HP_OK && i < p->hdr.frames; i++)
                p->index.push_back({get_le(t.data() + i * kIndexSize, 8), get_le(t.data() + i * kIndexSize + 8, 8)});
        }
    }
    if (st != HP_OK) {
        fclose(f);
        delete p;
        return st;
    }
    *out = p;
    return HP_OK;
}

hp_status hp_player_info(const hp_player* p, hp_animation_info* out) {
    if (!p || !out) return afail(HP_ERR_NULL, "player or out is NULL");
    out->width = p->hdr.w;
    out->height = p->hdr.h;
    out->depth = p->hdr.d;
    out->channels = p->hdr.c;
    out->frame_count = p->hdr.frames;
    out->dt = p->hdr.dt;
    out->compression = p->hdr.compression();
    out->delta_encoding = (p->hdr.flags >> 4) & 1;
    return HP_OK;
}

hp_status hp_player_read_frame(hp_player* p, uint64_t i, float* const* channels) {
    if (!p || !channels) return afail(HP_ERR_NULL, "player or channels is NULL");
    for (uint32_t c = 0; c < p->hdr.c; c++)
        if (!channels[c]) return afail(HP_ERR_NULL, "channel pointer is NULL");
    std::vector<uint8_t>* fr = nullptr;
    hp_status st = p->read(i, &fr);
    if (st != HP_OK) return st;
    const size_t nch = (size_t)p->hdr.w * p->hdr.h * p->hdr.d * 4;
    for (uint32_t c = 0; c < p->hdr.c; c++) memcpy(channels[c], fr->data() + c * nch, nch);
    return HP_OK;
}

hp_status hp_player_load_ctx(hp_player* p, uint64_t i, hp_ctx* ctx, float time, uint64_t step) {
    if (!p || !ctx) return afail(HP_ERR_NULL, "player or context is NULL");
    hp_config cfg{};
    hp_status st = hp_get_config(ctx, &cfg);
    if (st != HP_OK) return afail(st, "context config");
    if (cfg.width != p->hdr.w || cfg.height != p->hdr.h || cfg.depth != p->hdr.d || cfg.channels != p->hdr.c)
        return afail(HP_ERR_SHAPE, "the context's grid differs from the animation's");
    std::vector<uint8_t>* fr = nullptr;
    if ((st = p->read(i, &fr)) != HP_OK) return st;
    const size_t n = (size_t)p->hdr.w * p->hdr.h * p->hdr.d;
    std::vector<const float*> ptr(p->hdr.c);
    for (uint32_t c = 0; c < p->hdr.c; c++) ptr[c] = (const float*)(fr->data() + c * n * 4);
    st = hp_set_state(ctx, ptr.data(), time, step);
    if (st != HP_OK) return afail(st, hp_last_error(ctx));
    return HP_OK;
}

void hp_player_close(hp_player* p) {
    if (!p) return;
    if (p->f) fclose(p->f);
    delete p;
}

}  // extern "C"
