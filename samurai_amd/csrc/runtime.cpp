#include "runtime.hpp"

#include "host/parallel.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

namespace smr
{
    void hip_check(hipError_t e, const char* what)
    {
        if (e != hipSuccess)
        {
            throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
        }
    }

    // ------------------------------------------------------------------
    // timers
    // ------------------------------------------------------------------
    namespace
    {
        double now_s()
        {
            return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
        }
    }

    Timers& timers()
    {
        static Timers t;
        return t;
    }

    void Timers::add(const std::string& name, double seconds)
    {
        for (auto& e : entries)
        {
            if (e.first == name)
            {
                e.second.first += seconds;
                e.second.second += 1;
                return;
            }
        }
        entries.push_back({name, {seconds, 1}});
    }

    std::string Timers::report() const
    {
        std::string s = "{";
        for (size_t k = 0; k < entries.size(); ++k)
        {
            char buf[256];
            std::snprintf(buf, sizeof(buf), "%s\"%s\": {\"seconds\": %.6f, \"calls\": %lld}", k ? ", " : "", entries[k].first.c_str(),
                          entries[k].second.first, static_cast<long long>(entries[k].second.second));
            s += buf;
        }
        return s + "}";
    }

    ScopedTimer::ScopedTimer(Ctx* ctx, const char* name)
        : ctx_(ctx)
        , name_(name)
    {
        if (timers().enabled)
        {
            if (ctx_ && ctx_->has_device())
            {
                ctx_->sync();
            }
            t0_ = now_s();
        }
    }

    ScopedTimer::~ScopedTimer()
    {
        if (timers().enabled)
        {
            if (ctx_ && ctx_->has_device())
            {
                (void)hipStreamSynchronize(ctx_->stream);
            }
            timers().add(name_, now_s() - t0_);
        }
    }

    // ------------------------------------------------------------------
    // Blob
    // ------------------------------------------------------------------
    namespace
    {
        struct DevCache
        {
            std::mutex m;
            // block -> (size, stream it was allocated on)
            std::unordered_map<void*, std::pair<size_t, hipStream_t>> live;
            // per stream: free blocks by size, and their total
            std::unordered_map<hipStream_t, std::pair<std::multimap<size_t, void*>, size_t>> free;
        };

        DevCache& dev_cache()
        {
            // never destroyed: blocks may be freed by static destructors
            static DevCache* c = new DevCache;
            return *c;
        }

        // 256-byte multiples up to 4 KiB, then 8 size classes per power of two
        size_t block_size(size_t n)
        {
            n = std::max<size_t>(n, 256);
            if (n <= 4096)
            {
                return (n + 255) & ~size_t(255);
            }
            const int b       = 63 - __builtin_clzll(uint64_t(n));
            const size_t step = size_t(1) << (b - 3);
            return (n + step - 1) & ~(step - 1);
        }

        // frees the cached blocks of s, largest first, down to `keep` bytes (lock held)
        void trim_locked(DevCache& c, hipStream_t s, size_t keep)
        {
            auto it = c.free.find(s);
            if (it == c.free.end())
            {
                return;
            }
            auto& [blocks, total] = it->second;
            while (total > keep && !blocks.empty())
            {
                auto last = std::prev(blocks.end());
                (void)hipFreeAsync(last->second, s);
                total -= last->first;
                blocks.erase(last);
            }
        }
    } // namespace

    void* dev_alloc(size_t bytes, hipStream_t s)
    {
        const size_t n = block_size(bytes);
        DevCache& c    = dev_cache();
        std::lock_guard<std::mutex> lk(c.m);
        auto& [blocks, total] = c.free[s];
        // best fit, at most 1/4 larger than asked
        auto it = blocks.lower_bound(n);
        if (it != blocks.end() && it->first <= n + n / 4)
        {
            void* p = it->second;
            c.live.emplace(p, std::make_pair(it->first, s));
            total -= it->first;
            blocks.erase(it);
            return p;
        }
        void* p       = nullptr;
        hipError_t e  = hipMallocAsync(&p, n, s);
        if (e == hipErrorOutOfMemory && total > 0)
        {
            (void)hipGetLastError();
            trim_locked(c, s, 0);
            e = hipMallocAsync(&p, n, s);
        }
        hip_check(e, "hipMallocAsync");
        c.live.emplace(p, std::make_pair(n, s));
        return p;
    }

    void dev_free(void* p, hipStream_t s)
    {
        if (!p)
        {
            return;
        }
        DevCache& c = dev_cache();
        std::lock_guard<std::mutex> lk(c.m);
        auto it = c.live.find(p);
        if (it == c.live.end() || it->second.second != s)
        {
            // not from dev_alloc on s: no reuse is ordered after its last use
            if (it != c.live.end())
            {
                c.live.erase(it);
            }
            (void)hipFreeAsync(p, s);
            return;
        }
        const size_t n = it->second.first;
        c.live.erase(it);
        auto& [blocks, total] = c.free[s];
        blocks.emplace(n, p);
        total += n;
        if (total > kDevCacheLimit)
        {
            trim_locked(c, s, kDevCacheLimit / 2);
        }
    }

    void dev_cache_release(hipStream_t s)
    {
        DevCache& c = dev_cache();
        std::lock_guard<std::mutex> lk(c.m);
        trim_locked(c, s, 0);
        c.free.erase(s);
    }

    Blob::~Blob()
    {
        dev_free(d_, stream_);
    }

    size_t Blob::scratch(size_t n)
    {
        const size_t off = (scratch_ + 255) & ~size_t(255);
        scratch_         = off + std::max<size_t>(n, 8);
        return off;
    }

    size_t Blob::add(const void* src, size_t n)
    {
        const size_t off = (host_size_ + 255) & ~size_t(255);
        host_size_       = off + std::max<size_t>(n, 8);
        if (n)
        {
            parts_.push_back({src, n, off});
        }
        return off;
    }

    void Blob::write_parts(char* dst)
    {
        auto one = [&](int i)
        {
            const Part& p = parts_[size_t(i)];
            std::memcpy(dst + p.off, p.src, p.n);
        };
        if (host_size_ >= (size_t(1) << 20))
        {
            parallel_for(int(parts_.size()), one);
        }
        else
        {
            for (int i = 0; i < int(parts_.size()); ++i)
            {
                one(i);
            }
        }
        std::vector<Part>().swap(parts_);
        std::vector<std::shared_ptr<void>>().swap(keep_);
    }

    void Blob::upload(hipStream_t s)
    {
        if (d_)
        {
            throw std::logic_error("blob uploaded twice");
        }
        alloc(s);
        copy();
    }

    void Blob::alloc(hipStream_t s)
    {
        stream_ = s;
        d_      = static_cast<char*>(dev_alloc(bytes(), s));
    }

    void Blob::copy()
    {
        if (!host_size_)
        {
            return;
        }
        if (!staged_)
        {
            host_.reset(new char[host_size_]);
            write_parts(host_.get());
            staged_ = host_.get();
        }
        hip_check(hipMemcpyAsync(d_, staged_, host_size_, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync(blob)");
    }

    void Blob::stage(char* dst)
    {
        if (host_size_)
        {
            write_parts(dst);
        }
        staged_ = dst;
    }

    PinnedStaging::~PinnedStaging()
    {
        if (pending_)
        {
            (void)hipEventSynchronize(done_);
        }
        if (h_)
        {
            (void)hipHostFree(h_);
        }
        if (done_)
        {
            (void)hipEventDestroy(done_);
        }
    }

    char* PinnedStaging::acquire(size_t n)
    {
        if (pending_)
        {
            hip_check(hipEventSynchronize(done_), "hipEventSynchronize(staging)");
            pending_ = false;
        }
        if (n > cap_)
        {
            if (h_)
            {
                hip_check(hipHostFree(h_), "hipHostFree(staging)");
                h_ = nullptr;
            }
            cap_ = std::max(n, cap_ + cap_ / 2);
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_), cap_, hipHostMallocDefault), "hipHostMalloc(staging)");
        }
        return h_;
    }

    void PinnedStaging::release(hipStream_t s)
    {
        if (!done_)
        {
            hip_check(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "hipEventCreate(staging)");
        }
        hip_check(hipEventRecord(done_, s), "hipEventRecord(staging)");
        pending_ = true;
    }

    // ------------------------------------------------------------------
    // Ctx
    // ------------------------------------------------------------------
    Ctx::~Ctx()
    {
        // work still queued on the stream may use d_err and the cached blocks
        if (stream)
        {
            (void)hipStreamSynchronize(stream);
        }
        staging.reset();
        readback.reset();
        if (runs_back)
        {
            (void)hipEventDestroy(runs_back);
        }
        if (d_err)
        {
            (void)hipFree(d_err);
        }
        if (stream)
        {
            dev_cache_release(stream);
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }

    void Ctx::mark_runs_back()
    {
        if (!runs_back)
        {
            hip_check(hipEventCreateWithFlags(&runs_back, hipEventDisableTiming), "hipEventCreate(runs)");
        }
        hip_check(hipEventRecord(runs_back, stream), "hipEventRecord(runs)");
    }

    void Ctx::require_device() const
    {
        if (!has_device())
        {
            throw std::runtime_error("this context has no GPU: device kernels are unavailable");
        }
    }

    void Ctx::sync()
    {
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void Ctx::check_resolve()
    {
        int err = 0;
        hip_check(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(err)");
        sync();
        if (err)
        {
            hip_check(hipMemsetAsync(d_err, 0, sizeof(int), stream), "hipMemsetAsync(err)");
            throw std::out_of_range("a work-list row is missing from the reference interval table");
        }
    }

    // ------------------------------------------------------------------
    // device images
    // ------------------------------------------------------------------
    // ------------------------------------------------------------------
    // Packer: several arrays of one device image in one allocation
    // ------------------------------------------------------------------
    void Packer::table(const MRMeshData& m, dev::Table& t)
    {
        std::vector<int64_t> key, row_off(MAXL + 1, 0), rp, idx;
        std::vector<int32_t> rs, re;
        rp.push_back(0);
        for (int l = 0; l < MAXL; ++l)
        {
            row_off[size_t(l)] = int64_t(key.size());
            const LevelSet& r  = m.ids[REFERENCE][size_t(l)];
            for (size_t row = 0; row < r.nrows(); ++row)
            {
                key.push_back((int64_t(r.rz[row]) << 32) + (int64_t(r.ry[row]) + 2147483648LL));
                for (int64_t k = r.rp[row]; k < r.rp[row + 1]; ++k)
                {
                    rs.push_back(r.runs[size_t(k)].s);
                    re.push_back(r.runs[size_t(k)].e);
                    idx.push_back(m.ridx[size_t(l)][size_t(k)]);
                }
                rp.push_back(int64_t(rs.size()));
            }
        }
        row_off[MAXL] = int64_t(key.size());
        // the blob keeps these vectors until it is staged
        const size_t o_key = b_.add(std::move(key)), o_off = b_.add(std::move(row_off));
        const size_t o_rp = b_.add(std::move(rp)), o_rs = b_.add(std::move(rs));
        const size_t o_re = b_.add(std::move(re)), o_idx = b_.add(std::move(idx));
        Blob* b = &b_;
        binds_.push_back(
            [=, &t]
            {
                t.row_key = reinterpret_cast<const int64_t*>(b->dptr(o_key));
                t.row_off = reinterpret_cast<const int64_t*>(b->dptr(o_off));
                t.rp      = reinterpret_cast<const int64_t*>(b->dptr(o_rp));
                t.run_s   = reinterpret_cast<const int32_t*>(b->dptr(o_rs));
                t.run_e   = reinterpret_cast<const int32_t*>(b->dptr(o_re));
                t.run_idx = reinterpret_cast<const int64_t*>(b->dptr(o_idx));
            });
    }

    void Packer::recipes(DevRecipes& r)
    {
        const size_t od  = b_.add(r.host.dst.data(), r.host.dst.size() * 8);
        const size_t oo  = b_.add(r.host.op.data(), r.host.op.size() * 4);
        const size_t osp = b_.add(r.host.src_ptr.data(), r.host.src_ptr.size() * 8);
        const size_t os  = b_.add(r.host.src.data(), r.host.src.size() * 8);
        Blob* b          = &b_;
        binds_.push_back(
            [=, &r]
            {
                r.dst = reinterpret_cast<const int64_t*>(b->dptr(od));
                r.op  = reinterpret_cast<const int32_t*>(b->dptr(oo));
                r.sp  = reinterpret_cast<const int64_t*>(b->dptr(osp));
                r.src = reinterpret_cast<const int64_t*>(b->dptr(os));
            });
    }

    // the host arrays of a work list besides its runs: chunk hints (into l),
    // taps, 64-cell tiles and column walks (thread-safe across lists)
    Packer::ListPrep Packer::prepare_list(DevWorkList& l)
    {
        const WorkList& w     = l.host;
        const int64_t nruns   = w.nruns();
        const int64_t ncells  = w.ncells();
        const int64_t nchunks = (ncells + (1 << dev::HINT_SHIFT) - 1) >> dev::HINT_SHIFT;
        l.hint.assign(size_t(nchunks) + 1, 0);
        int64_t r = 0;
        for (int64_t j = 0; j < nchunks; ++j)
        {
            const int64_t c = j << dev::HINT_SHIFT;
            while (w.pref[size_t(r + 1)] <= c)
            {
                ++r;
            }
            l.hint[size_t(j)] = int32_t(r);
        }
        l.hint[size_t(nchunks)] = int32_t(std::max<int64_t>(nruns - 1, 0));
        std::vector<int8_t> taps;
        for (const Tap& t : w.taps)
        {
            taps.insert(taps.end(), {t.table, t.dl, t.dx, t.dy, t.dz, t.wlo, t.whi});
        }
        // 64-cell wave tiles of every run
        std::vector<int32_t> tiles;
        tiles.reserve(size_t(nruns) * 2 + size_t(ncells / 64) * 2);
        for (int64_t k = 0; k < nruns; ++k)
        {
            for (int32_t x = w.iv[size_t(5 * k)]; x < w.iv[size_t(5 * k + 1)]; x += 64)
            {
                tiles.push_back(int32_t(k));
                tiles.push_back(x);
            }
        }
        // column walks: chains of runs stacked in y with equal x ranges, cut in
        // 64-column strips of at most kWalkRows runs
        std::vector<int32_t> next, walks;
        int64_t nrows_total = 0;
        int32_t walk_stride = 0;
        if (w.walk)
        {
            // rows per walk: n + 2 row bases and 2 slots of walk data fit one
            // wavefront (n <= 60); a wave's rows are open DRAM pages at once, so
            // short walks stream better, at the price of the two extra rows
            // (10 measured best on the uniform L12 mesh, 16 about 1 % slower)
            static const int kWalkRows = []
            {
                const char* e = std::getenv("SMR_WALK_ROWS");
                const int v   = e ? std::atoi(e) : 10;
                return std::max(1, std::min(v, 60));
            }();
            walk_stride = kWalkRows + 4;
            next.assign(size_t(nruns), -1);
            std::vector<uint8_t> below(size_t(nruns), 0);
            // runs are sorted by level, then (z, y), then x: the run above r with
            // the same x range is found by a merge over consecutive rows
            int64_t r = 0;
            while (r < nruns)
            {
                const int32_t* a = &w.iv[size_t(5 * r)];
                int64_t e        = r; // end of the row of r
                while (e < nruns && w.iv[size_t(5 * e + 2)] == a[2] && w.iv[size_t(5 * e + 3)] == a[3] && w.iv[size_t(5 * e + 4)] == a[4])
                {
                    ++e;
                }
                // row above: starts at e if it is row y+1 of the same level and z
                if (e < nruns && w.iv[size_t(5 * e + 2)] == a[2] + 1 && w.iv[size_t(5 * e + 3)] == a[3] && w.iv[size_t(5 * e + 4)] == a[4])
                {
                    int64_t q = e;
                    for (int64_t k = r; k < e; ++k)
                    {
                        const int32_t x0 = w.iv[size_t(5 * k)], x1 = w.iv[size_t(5 * k + 1)];
                        while (q < nruns && w.iv[size_t(5 * q + 2)] == a[2] + 1 && w.iv[size_t(5 * q + 3)] == a[3] &&
                               w.iv[size_t(5 * q + 4)] == a[4] && w.iv[size_t(5 * q)] < x0)
                        {
                            ++q;
                        }
                        if (q < nruns && w.iv[size_t(5 * q + 2)] == a[2] + 1 && w.iv[size_t(5 * q + 3)] == a[3] &&
                            w.iv[size_t(5 * q + 4)] == a[4] && w.iv[size_t(5 * q)] == x0 && w.iv[size_t(5 * q + 1)] == x1)
                        {
                            next[size_t(k)] = int32_t(q);
                            below[size_t(q)] = 1;
                        }
                    }
                }
                r = e;
            }
            for (int64_t s0 = 0; s0 < nruns; ++s0)
            {
                if (below[size_t(s0)])
                {
                    continue;
                }
                // a band of up to kWalkRows runs, then every 64-column strip of it:
                // neighbouring strips (which share the x-1 / x+1 cache lines) are
                // consecutive walks, run by neighbouring waves of one XCD
                const int32_t x0 = w.iv[size_t(5 * s0)], x1 = w.iv[size_t(5 * s0 + 1)];
                int64_t k = s0;
                while (k >= 0)
                {
                    const int64_t first = k;
                    int32_t n           = 0;
                    while (k >= 0 && n < kWalkRows)
                    {
                        ++n;
                        k = next[size_t(k)];
                    }
                    for (int32_t xs = x0; xs < x1; xs += 64)
                    {
                        walks.insert(walks.end(), {int32_t(first), xs, n, int32_t(nrows_total)});
                        nrows_total += walk_stride;
                    }
                }
            }
        }
        ListPrep prep;
        prep.taps        = std::move(taps);
        prep.tiles       = std::move(tiles);
        prep.next        = std::move(next);
        prep.walks       = std::move(walks);
        prep.nrows_total = nrows_total;
        prep.walk_stride = walk_stride;
        return prep;
    }

    void Packer::commit_list(DevWorkList& l, ListPrep&& prep)
    {
        const WorkList& w    = l.host;
        const int64_t nruns  = w.nruns();
        const int64_t ncells = w.ncells();
        std::vector<int8_t>& taps    = prep.taps;
        std::vector<int32_t>& tiles  = prep.tiles;
        std::vector<int32_t>& next   = prep.next;
        std::vector<int32_t>& walks  = prep.walks;
        const int64_t nrows_total    = prep.nrows_total;
        const int32_t walk_stride    = prep.walk_stride;
        const size_t oiv  = b_.add(w.iv.data(), w.iv.size() * 4);
        const size_t opf  = b_.add(w.pref.data(), w.pref.size() * 8);
        const size_t oh   = b_.add(l.hint.data(), l.hint.size() * 4);
        const int64_t ntiles = int64_t(tiles.size() / 2);
        const int64_t nwalks = int64_t(walks.size() / 4);
        // the blob keeps the vectors built here until it is staged
        const size_t ot   = b_.add(std::move(taps));
        const size_t otl  = b_.add(std::move(tiles));
        const size_t onx  = b_.add(std::move(next));
        const size_t owk  = b_.add(std::move(walks));
        const size_t owr  = b_.scratch(size_t(nrows_total) * 8);
        const size_t obas = b_.scratch(size_t(nruns) * w.taps.size() * 8);
        const size_t ocr  = b_.scratch(size_t(ncells) * 4);
        const int ntaps      = int(w.taps.size());
        Blob* b              = &b_;
        binds_.push_back(
            [=, &l]
            {
                l.d.iv     = reinterpret_cast<const int32_t*>(b->dptr(oiv));
                l.d.pref   = reinterpret_cast<const int64_t*>(b->dptr(opf));
                l.d.hint   = reinterpret_cast<const int32_t*>(b->dptr(oh));
                l.d.taps   = reinterpret_cast<const int8_t*>(b->dptr(ot));
                l.d.tiles  = reinterpret_cast<const int32_t*>(b->dptr(otl));
                l.d.next   = nwalks > 0 ? reinterpret_cast<const int32_t*>(b->dptr(onx)) : nullptr;
                l.d.walks  = nwalks > 0 ? reinterpret_cast<const int32_t*>(b->dptr(owk)) : nullptr;
                l.d.walk_rows = nwalks > 0 ? reinterpret_cast<int64_t*>(b->sptr(owr)) : nullptr;
                l.d.walk_stride = walk_stride;
                l.d.nwalks = nwalks;
                l.d.base   = reinterpret_cast<int64_t*>(b->sptr(obas));
                l.d.cellrun = ncells > 0 ? reinterpret_cast<int32_t*>(b->sptr(ocr)) : nullptr;
                l.d.ntiles = ntiles;
                l.d.nruns  = nruns;
                l.d.ncells = ncells;
                l.d.ntaps  = ntaps;
            });
    }

    void Packer::list(DevWorkList& l)
    {
        commit_list(l, prepare_list(l));
    }

    void Packer::device_list(DevWorkList& l, const int32_t* iv, const int64_t* pref, const int32_t* hint, int64_t nruns, int64_t ncells)
    {
        std::vector<int8_t> taps;
        for (const Tap& t : l.host.taps)
        {
            taps.insert(taps.end(), {t.table, t.dl, t.dx, t.dy, t.dz, t.wlo, t.whi});
        }
        const int ntaps   = int(l.host.taps.size());
        const size_t ot   = b_.add(std::move(taps));
        const size_t obas = b_.scratch(size_t(nruns) * size_t(ntaps) * 8);
        const size_t ocr  = b_.scratch(size_t(ncells) * 4);
        Blob* b           = &b_;
        binds_.push_back(
            [=, &l]
            {
                l.d         = dev::List{};
                l.d.iv      = iv;
                l.d.pref    = pref;
                l.d.hint    = hint;
                l.d.taps    = reinterpret_cast<const int8_t*>(b->dptr(ot));
                l.d.base    = reinterpret_cast<int64_t*>(b->sptr(obas));
                l.d.cellrun = ncells > 0 ? reinterpret_cast<int32_t*>(b->sptr(ocr)) : nullptr;
                l.d.nruns   = nruns;
                l.d.ncells  = ncells;
                l.d.ntaps   = ntaps;
            });
    }

    void Packer::lists(const std::vector<DevWorkList*>& ls)
    {
        std::vector<ListPrep> preps(ls.size());
        parallel_for(int(ls.size()), [&](int i) { preps[size_t(i)] = prepare_list(*ls[size_t(i)]); });
        for (size_t i = 0; i < ls.size(); ++i)
        {
            commit_list(*ls[i], std::move(preps[i]));
        }
    }

    void Packer::finish(hipStream_t s)
    {
        b_.alloc(s);
        for (auto& f : binds_)
        {
            f();
        }
        binds_.clear();
        b_.copy();
    }

    namespace
    {
        void upload_recipes(Ctx& ctx, DevRecipes& r)
        {
            r.blob = std::make_unique<Blob>();
            Packer p(*r.blob);
            p.recipes(r);
            p.finish(ctx.stream);
        }
    } // namespace

    void DevWorkList::cells_of_levels(int l0, int l1, int64_t& c0, int64_t& c1) const
    {
        if (device_built)
        {
            l0 = std::max(l0, 0);
            l1 = std::min(l1, MAXL);
            c0 = c1 = 0;
            if (l1 > l0)
            {
                c0 = level_cell[size_t(l0)];
                c1 = level_cell[size_t(l1)];
            }
            return;
        }
        int64_t r0    = 0;
        int64_t nr    = host.runs_of_levels(l0, l1, r0);
        c0            = host.pref[size_t(r0)];
        c1            = host.pref[size_t(r0 + nr)];
    }

    // the blob's host arrays through the context's pinned staging, then its
    // allocation, pointer binds and copy
    void finish_staged(Ctx& ctx, Packer& p, Blob& b)
    {
        b.stage(ctx.staging->acquire(b.host_bytes()));
        p.finish(ctx.stream);
        ctx.staging->release(ctx.stream);
    }

    void upload_list(Ctx& ctx, DevWorkList& l, const dev::Table& t0, const dev::Table& t1)
    {
        l.blob = std::make_unique<Blob>();
        Packer p(*l.blob);
        p.list(l);
        finish_staged(ctx, p, *l.blob);
        dev::resolve(l.d, t0, t1, ctx.d_err, ctx.stream);
        dev::walk_rows(l.d, ctx.stream);
    }

    std::unique_ptr<DevMesh> make_dev_mesh(Ctx& ctx, const MRMeshData& m)
    {
        ctx.require_device();
        auto dm        = std::make_unique<DevMesh>();
        dm->table_blob = std::make_unique<Blob>();
        Packer p(*dm->table_blob);
        p.table(m, dm->table);
        dm->cells.host = build_cells_list(m);
        p.list(dm->cells);
        finish_staged(ctx, p, *dm->table_blob);
        dev::resolve(dm->cells.d, dm->table, dm->table, ctx.d_err, ctx.stream);
        dev::walk_rows(dm->cells.d, ctx.stream);
        ctx.check_resolve();
        return dm;
    }

    HostImage build_host_image(const MRMeshData& m, const MRMeshData* old)
    {
        HostImage h;
        build_programs(m, h.cells, h.ghost, h.adapt, old, old ? h.uf : nullptr);
        h.has_uf = old != nullptr;
        return h;
    }

    void pack_image(Packer& p, DevMesh& dm, const MRMeshData& m, HostImage&& h)
    {
        p.table(m, dm.table);
        dm.outer.resize(h.ghost.outer.size());
        for (size_t l = 0; l < h.ghost.outer.size(); ++l)
        {
            dm.outer[l].host = std::move(h.ghost.outer[l]);
            p.recipes(dm.outer[l]);
        }
        dm.cells.host      = std::move(h.cells);
        dm.projection.host = std::move(h.ghost.projection);
        dm.prediction.host = std::move(h.ghost.prediction);
        dm.detail.host     = std::move(h.adapt.detail);
        dm.criteria.host   = std::move(h.adapt.criteria);
        dm.maximum.host    = std::move(h.adapt.maximum);
        std::vector<DevWorkList*> ls = dm.lists();
        if (h.has_uf)
        {
            for (int k = 0; k < 3; ++k)
            {
                dm.uf[k].host = std::move(h.uf[k]);
                ls.push_back(&dm.uf[k]);
            }
            dm.has_uf = true;
        }
        p.lists(ls);
        dm.has_ghost = dm.has_adapt = true;
    }

    std::unique_ptr<DevMesh> upload_image(Ctx& ctx, const MRMeshData& m, HostImage&& h, bool check, const dev::Table* old_table)
    {
        if (h.has_uf && !old_table)
        {
            throw std::logic_error("update_fields lists need the table of the mesh they come from");
        }
        ctx.require_device();
        auto dm        = std::make_unique<DevMesh>();
        dm->table_blob = std::make_unique<Blob>();
        Packer p(*dm->table_blob);
        pack_image(p, *dm, m, std::move(h));
        finish_staged(ctx, p, *dm->table_blob);
        // every list of the image (and its update_fields lists) in one launch
        dev::ResolveBatch rb{};
        rb.t_new = dm->table;
        rb.t_old = old_table ? *old_table : dm->table;
        rb.n     = 0;
        auto add = [&](const DevWorkList& l, bool old)
        {
            if (rb.n >= dev::kResolveJobs)
            {
                throw std::logic_error("resolve batch: too many lists");
            }
            rb.lists[rb.n] = l.d;
            rb.old[rb.n]   = old ? 1 : 0;
            rb.start[rb.n + 1] = rb.start[rb.n] + dev::resolve_range(l.d);
            ++rb.n;
        };
        for (DevWorkList* l : dm->lists())
        {
            add(*l, false);
        }
        if (dm->has_uf)
        {
            for (const DevWorkList& l : dm->uf)
            {
                add(l, true);
            }
        }
        dev::resolve_batch(rb, ctx.d_err, ctx.stream);
        dev::walk_rows(dm->cells.d, ctx.stream);
        if (check)
        {
            ctx.check_resolve();
        }
        return dm;
    }

    namespace
    {
        size_t align256(size_t n)
        {
            return (n + 255) & ~size_t(255);
        }

        int64_t list_chunks(int64_t ncells)
        {
            return (ncells + (int64_t(1) << dev::HINT_SHIFT) - 1) >> dev::HINT_SHIFT;
        }
    } // namespace

    size_t list_from_table(const uint64_t* t, DeviceLists::One& o, dev::ListOut& w)
    {
        auto runs_of  = [](uint64_t v) { return int64_t(v >> 32); };
        auto cells_of = [](uint64_t v) { return int64_t(v & 0xFFFFFFFFull); };
        o.nruns       = runs_of(t[dev::SM_SEG - 1]) - runs_of(t[0]);
        o.ncells      = cells_of(t[dev::SM_SEG - 1]) - cells_of(t[0]);
        o.level_run.resize(dev::SM_SEG);
        o.level_cell.resize(dev::SM_SEG);
        for (int l = 0; l < dev::SM_SEG; ++l)
        {
            o.level_run[size_t(l)]  = runs_of(t[l]) - runs_of(t[0]);
            o.level_cell[size_t(l)] = cells_of(t[l]) - cells_of(t[0]);
            if (l > 0 && (o.level_run[size_t(l)] < o.level_run[size_t(l - 1)] || o.level_cell[size_t(l)] < o.level_cell[size_t(l - 1)]))
            {
                throw std::logic_error("device work lists: inconsistent counts at level " + std::to_string(l));
            }
        }
        w.run0   = runs_of(t[0]);
        w.cell0  = cells_of(t[0]);
        w.nruns  = o.nruns;
        w.ncells = o.ncells;
        return align256(size_t(std::max<int64_t>(o.nruns, 1)) * 20) + align256(size_t(o.nruns + 1) * 8) +
               align256(size_t(list_chunks(o.ncells) + 1) * 4);
    }

    void bind_list_arrays(char* d, DeviceLists::One& o, dev::ListOut& w)
    {
        const size_t a = align256(size_t(std::max<int64_t>(o.nruns, 1)) * 20), b = align256(size_t(o.nruns + 1) * 8);
        w.iv           = reinterpret_cast<int32_t*>(d);
        w.pref         = reinterpret_cast<int64_t*>(d + a);
        w.hint         = reinterpret_cast<int32_t*>(d + a + b);
        o.iv           = w.iv;
        o.pref         = w.pref;
        o.hint         = w.hint;
    }

    void pack_device_list(Packer& p, DevWorkList& w, int kind, const MeshConfig& cfg, const DeviceLists::One& o)
    {
        w.host           = WorkList();
        w.host.taps      = list_taps(kind, cfg.dim, cfg.pred_s);
        w.host.level_run = o.level_run;
        w.level_cell     = o.level_cell;
        w.device_built   = true;
        p.device_list(w, o.iv, o.pref, o.hint, o.nruns, o.ncells);
    }

    void pack_device_lists(Packer& p, DevMesh& dm, const MRMeshData& m, std::vector<Recipes>&& outer, WorkList&& cells, const DeviceLists::One* l,
                           std::shared_ptr<DevBuffer> mem)
    {
        p.table(m, dm.table);
        dm.outer.resize(outer.size());
        for (size_t k = 0; k < outer.size(); ++k)
        {
            dm.outer[k].host = std::move(outer[k]);
            p.recipes(dm.outer[k]);
        }
        dm.cells.host = std::move(cells);
        p.list(dm.cells);
        DevWorkList* dls[dev::DL_UF_COPY] = {&dm.projection, &dm.prediction, &dm.detail, &dm.criteria, &dm.maximum};
        for (int k = 0; k < dev::DL_UF_COPY; ++k)
        {
            pack_device_list(p, *dls[k], k, m.cfg, l[k]);
        }
        dm.list_mem  = std::move(mem);
        dm.has_ghost = dm.has_adapt = true;
    }

    std::unique_ptr<DevMesh> upload_image_device_lists(Ctx& ctx, const MRMeshData& m, std::vector<Recipes>&& outer, WorkList&& cells,
                                                       DeviceLists&& dl, const dev::Table* old_table)
    {
        const bool uf = dl.nlists > dev::DL_UF_COPY;
        if (uf && !old_table)
        {
            throw std::logic_error("update_fields lists need the table of the mesh they come from");
        }
        ctx.require_device();
        auto dm        = std::make_unique<DevMesh>();
        dm->table_blob = std::make_unique<Blob>();
        Packer p(*dm->table_blob);
        pack_device_lists(p, *dm, m, std::move(outer), std::move(cells), dl.l, dl.mem);
        if (uf)
        {
            for (int k = 0; k < 3; ++k)
            {
                pack_device_list(p, dm->uf[k], dev::DL_UF_COPY + k, m.cfg, dl.l[dev::DL_UF_COPY + k]);
            }
        }
        dm->mesh_sets     = std::move(dl.sets);
        dm->has_mesh_sets = dm->mesh_sets.p != nullptr;
        dm->has_uf        = uf;
        finish_staged(ctx, p, *dm->table_blob);
        // every list of the image (and its update_fields lists) in one launch
        dev::ResolveBatch rb{};
        rb.t_new = dm->table;
        rb.t_old = old_table ? *old_table : dm->table;
        rb.n     = 0;
        auto add = [&](const DevWorkList& l, bool old)
        {
            rb.lists[rb.n]     = l.d;
            rb.old[rb.n]       = old ? 1 : 0;
            rb.start[rb.n + 1] = rb.start[rb.n] + dev::resolve_range(l.d);
            ++rb.n;
        };
        for (DevWorkList* l : dm->lists())
        {
            add(*l, false);
        }
        if (uf)
        {
            for (const DevWorkList& l : dm->uf)
            {
                add(l, true);
            }
        }
        dev::resolve_batch(rb, ctx.d_err, ctx.stream);
        dev::walk_rows(dm->cells.d, ctx.stream);
        return dm;
    }

    int check_device_lists(Mesh& m)
    {
        if (!m.dev || !m.dev->projection.device_built)
        {
            return 1;
        }
        DevMesh& dm = *m.dev;
        WorkList cells;
        GhostProgram gp;
        AdaptProgram ap;
        build_programs(m.host, cells, gp, ap);
        const char* names[5] = {"projection", "prediction", "detail", "criteria", "maximum"};
        DevWorkList* d[5]    = {&dm.projection, &dm.prediction, &dm.detail, &dm.criteria, &dm.maximum};
        const WorkList* h[5] = {&gp.projection, &gp.prediction, &ap.detail, &ap.criteria, &ap.maximum};
        for (int k = 0; k < 5; ++k)
        {
            const WorkList& w = *h[k];
            const dev::List& l = d[k]->d;
            auto fail = [&](const char* what) { throw std::logic_error(std::string("device-built ") + names[k] + " list: " + what + " differs"); };
            if (l.nruns != w.nruns() || l.ncells != w.ncells())
            {
                fail("size");
            }
            if (d[k]->host.taps.size() != w.taps.size())
            {
                fail("taps");
            }
            std::vector<int32_t> iv(w.iv.size());
            std::vector<int64_t> pref(w.pref.size());
            const int64_t nchunks = (w.ncells() + (int64_t(1) << dev::HINT_SHIFT) - 1) >> dev::HINT_SHIFT;
            std::vector<int32_t> hint(size_t(nchunks) + 1);
            hip_check(hipMemcpyAsync(iv.data(), l.iv, iv.size() * 4, hipMemcpyDeviceToHost, m.ctx->stream), "hipMemcpyAsync(check)");
            hip_check(hipMemcpyAsync(pref.data(), l.pref, pref.size() * 8, hipMemcpyDeviceToHost, m.ctx->stream), "hipMemcpyAsync(check)");
            hip_check(hipMemcpyAsync(hint.data(), l.hint, hint.size() * 4, hipMemcpyDeviceToHost, m.ctx->stream), "hipMemcpyAsync(check)");
            m.ctx->sync();
            if (iv != w.iv)
            {
                fail("runs");
            }
            if (pref != w.pref)
            {
                fail("cell prefix");
            }
            int64_t r = 0;
            for (int64_t j = 0; j < nchunks; ++j)
            {
                while (w.pref[size_t(r + 1)] <= (j << dev::HINT_SHIFT))
                {
                    ++r;
                }
                if (hint[size_t(j)] != r)
                {
                    fail("hint");
                }
            }
            if (hint[size_t(nchunks)] != int32_t(std::max<int64_t>(w.nruns() - 1, 0)))
            {
                fail("last hint");
            }
            for (int l0 = 0; l0 < MAXL; ++l0)
            {
                int64_t a, b, r0;
                d[k]->cells_of_levels(l0, l0 + 1, a, b);
                const int64_t nr = w.runs_of_levels(l0, l0 + 1, r0);
                if (a != w.pref[size_t(r0)] || b != w.pref[size_t(r0 + nr)] || d[k]->host.level_run[size_t(l0)] != w.level_run[size_t(l0)])
                {
                    fail("level offsets");
                }
            }
        }
        return 0;
    }

    void Mesh::ensure_host_lists()
    {
        DevMesh& dm = device_adapt();
        if (!dm.projection.device_built && !dm.detail.device_built)
        {
            return;
        }
        WorkList cells;
        GhostProgram gp;
        AdaptProgram ap;
        build_programs(host, cells, gp, ap);
        DevWorkList* d[5]    = {&dm.projection, &dm.prediction, &dm.detail, &dm.criteria, &dm.maximum};
        WorkList* h[5]       = {&gp.projection, &gp.prediction, &ap.detail, &ap.criteria, &ap.maximum};
        for (int k = 0; k < 5; ++k)
        {
            if (h[k]->nruns() != d[k]->d.nruns || h[k]->ncells() != d[k]->d.ncells)
            {
                throw std::logic_error("device-built work list differs from the host's");
            }
            d[k]->host         = std::move(*h[k]);
            d[k]->device_built = false;
            d[k]->level_cell.clear();
        }
    }

    DevMesh::DevMesh()
    {
        static std::atomic<uint64_t> next{1};
        serial = next++;
    }

    DevMesh& Mesh::device()
    {
        if (!dev)
        {
            dev = make_dev_mesh(*ctx, host);
        }
        return *dev;
    }

    DevMesh& Mesh::device_ghost()
    {
        DevMesh& dm = device();
        if (!dm.has_ghost)
        {
            ScopedTimer timer(ctx, "ghost program build");
            GhostProgram gp = build_ghost_program(host);
            dm.outer.resize(gp.outer.size());
            for (size_t l = 0; l < gp.outer.size(); ++l)
            {
                dm.outer[l].host = std::move(gp.outer[l]);
                upload_recipes(*ctx, dm.outer[l]);
            }
            dm.projection.host = std::move(gp.projection);
            dm.prediction.host = std::move(gp.prediction);
            upload_list(*ctx, dm.projection, dm.table, dm.table);
            upload_list(*ctx, dm.prediction, dm.table, dm.table);
            ctx->check_resolve();
            dm.has_ghost = true;
        }
        return dm;
    }

    DevMesh& Mesh::device_adapt()
    {
        DevMesh& dm = device_ghost();
        if (!dm.has_adapt)
        {
            ScopedTimer timer(ctx, "adapt program build");
            AdaptProgram ap  = build_adapt_program(host);
            dm.detail.host   = std::move(ap.detail);
            dm.criteria.host = std::move(ap.criteria);
            dm.maximum.host  = std::move(ap.maximum);
            upload_list(*ctx, dm.detail, dm.table, dm.table);
            upload_list(*ctx, dm.criteria, dm.table, dm.table);
            upload_list(*ctx, dm.maximum, dm.table, dm.table);
            ctx->check_resolve();
            dm.has_adapt = true;
        }
        return dm;
    }

    void Mesh::replace(MRMeshData&& m, std::unique_ptr<DevMesh>&& d)
    {
        host = std::move(m);
        dev  = std::move(d);
        ++version;
    }

    // ------------------------------------------------------------------
    // Field
    // ------------------------------------------------------------------
    Field::~Field()
    {
        if (d)
        {
            dev_free(d, mesh->ctx->stream);
        }
    }

    void Field::allocate(bool zero)
    {
        mesh->ctx->require_device();
        if (d)
        {
            dev_free(d, mesh->ctx->stream);
            d = nullptr;
        }
        size = mesh->host.nref;
        d    = static_cast<double*>(dev_alloc(size_t(std::max<int64_t>(size, 1)) * sizeof(double), mesh->ctx->stream));
        if (zero)
        {
            hip_check(hipMemsetAsync(d, 0, size_t(size) * sizeof(double), mesh->ctx->stream), "hipMemsetAsync(field)");
        }
        version = mesh->version;
    }

    void Field::check_current() const
    {
        if (version != mesh->version || !d)
        {
            throw std::runtime_error("field '" + name + "' was not resized after its mesh changed");
        }
    }

    AdaptState::~AdaptState()
    {
        if (rel_max)
        {
            (void)hipFree(rel_max);
        }
        if (detail)
        {
            (void)hipFree(detail);
        }
        if (tag)
        {
            (void)hipFree(tag);
        }
        if (host_tag)
        {
            (void)hipHostFree(host_tag);
        }
    }

    void AdaptState::copy_last(uint8_t* tags, double* details, hipStream_t s) const
    {
        if (tags && last_size > 0)
        {
            if (!host_tag_valid)
            {
                hip_check(hipMemcpyAsync(tags, tag, size_t(last_size), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(tag)");
                hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
            }
            else
            {
                std::memcpy(tags, host_tag, size_t(last_size));
            }
        }
        if (details && last_size > 0)
        {
            hip_check(hipMemcpyAsync(details, detail, size_t(last_size) * size_t(ncomp()) * 8, hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync(detail)");
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        }
    }

    // ------------------------------------------------------------------
    // update_ghost_mr (algorithm/update_ghost_mr.hpp:194-237)
    // ------------------------------------------------------------------
    void SeqPlan::build(std::vector<dev::SeqItem>&& all, hipStream_t s)
    {
        items = std::move(all);
        steps.clear();
        for (int k = 0; k < int(items.size());)
        {
            if (items[size_t(k)].c1 - items[size_t(k)].c0 > kSeqSmall)
            {
                steps.push_back({false, k, k + 1});
                ++k;
                continue;
            }
            int e = k + 1;
            while (e < int(items.size()) && items[size_t(e)].c1 - items[size_t(e)].c0 <= kSeqSmall)
            {
                ++e;
            }
            steps.push_back({true, k, e});
            k = e;
        }
        if (!items.empty())
        {
            void* d = d_items.get(items.size() * sizeof(dev::SeqItem), s);
            hip_check(hipMemcpyAsync(d, items.data(), items.size() * sizeof(dev::SeqItem), hipMemcpyHostToDevice, s), "hipMemcpyAsync(plan)");
        }
        built = true;
    }

    namespace
    {
        dev::SeqItem seq_item(int32_t op, int64_t c0, int64_t c1)
        {
            dev::SeqItem it{};
            it.op = op;
            it.c0 = c0;
            it.c1 = c1;
            return it;
        }

        // update_ghost_mr's stages in order (algorithm/update_ghost_mr.hpp:
        // 194-237): per level from the finest, the outer-ghost recipe stages
        // then the projection into the level below; then the predictions from
        // the coarsest level up
        // (for a field with a B.C., the stages of copies meant for B.C.-less
        // fields are left out)
        SeqPlan& ghost_plan(Ctx& ctx, DevMesh& dm, const MeshConfig& cfg, bool has_bc)
        {
            SeqPlan& plan = has_bc ? dm.ghost_plan : dm.ghost_plan_nobc;
            if (!plan.built)
            {
                std::vector<dev::SeqItem> items;
                for (int level = cfg.max_level; level >= 0; --level)
                {
                    const DevRecipes& r = dm.outer[size_t(level)];
                    const auto& sp      = r.host.stage_ptr;
                    for (size_t k = 0; k + 1 < sp.size(); ++k)
                    {
                        if (sp[k + 1] > sp[k] && !(has_bc && r.host.stage_without_bc_only(k)))
                        {
                            dev::SeqItem it = seq_item(dev::SEQ_RECIPES, sp[k], sp[k + 1]);
                            it.dst          = r.dst;
                            it.rop          = r.op;
                            it.sp           = r.sp;
                            it.src          = r.src;
                            items.push_back(it);
                        }
                    }
                    if (level > 0)
                    {
                        int64_t c0, c1;
                        dm.projection.cells_of_levels(level - 1, level, c0, c1);
                        if (c1 > c0)
                        {
                            items.push_back(seq_item(dev::SEQ_PROJECT, c0, c1));
                        }
                    }
                }
                for (int level = 1; level <= cfg.max_level; ++level)
                {
                    int64_t c0, c1;
                    dm.prediction.cells_of_levels(level, level + 1, c0, c1);
                    if (c1 > c0)
                    {
                        items.push_back(seq_item(dev::SEQ_PREDICT, c0, c1));
                    }
                }
                plan.build(std::move(items), ctx.stream);
            }
            return plan;
        }

        // maximum_op from the finest level down (mr/adapt.hpp:348-358)
        SeqPlan& max_plan(Ctx& ctx, DevMesh& dm, const MeshConfig& cfg)
        {
            if (!dm.max_plan.built)
            {
                std::vector<dev::SeqItem> items;
                for (int level = cfg.max_level; level > 0; --level)
                {
                    int64_t c0, c1;
                    dm.maximum.cells_of_levels(level - 1, level, c0, c1);
                    if (c1 > c0)
                    {
                        items.push_back(seq_item(dev::SEQ_MAXIMUM, c0, c1));
                    }
                }
                dm.max_plan.build(std::move(items), ctx.stream);
            }
            return dm.max_plan;
        }
    } // namespace

    void update_ghost_mr(Field& f)
    {
        f.check_current();
        Mesh& m         = *f.mesh;
        DevMesh& dm     = m.device_ghost();
        ScopedTimer timer(m.ctx, "ghost update");
        hipStream_t s   = m.ctx->stream;
        const int dim   = m.host.cfg.dim;
        const int ps    = m.host.cfg.pred_s;
        const double bc = f.bc_value;
        const SeqPlan& plan = ghost_plan(*m.ctx, dm, m.host.cfg, f.has_bc);
        const auto* d_items = static_cast<const dev::SeqItem*>(plan.d_items.p);
        for (const SeqPlan::Step& st : plan.steps)
        {
            if (st.seq)
            {
                dev::sequence(dim, ps, d_items, st.i0, st.i1, dm.projection.d, dm.prediction.d, dev::List{}, f.d, nullptr, f.has_bc, bc, s);
                continue;
            }
            const dev::SeqItem& it = plan.items[size_t(st.i0)];
            switch (it.op)
            {
                case dev::SEQ_RECIPES: dev::recipes(it.dst, it.rop, it.sp, it.src, it.c0, it.c1, f.d, f.has_bc, bc, s); break;
                case dev::SEQ_PROJECT: dev::project(dim, dm.projection.d, it.c0, it.c1, f.d, f.d, s); break;
                case dev::SEQ_PREDICT: dev::predict_fine(dim, ps, dm.prediction.d, it.c0, it.c1, f.d, s); break;
                default: break;
            }
        }
    }

    void upwind(const double* a, double dt, Field& u, Field& unp1)
    {
        u.check_current();
        unp1.check_current();
        if (u.mesh != unp1.mesh)
        {
            throw std::invalid_argument("upwind: fields on different meshes");
        }
        Mesh& m     = *u.mesh;
        DevMesh& dm = m.device();
        ScopedTimer timer(m.ctx, "field expressions");
        const int dim = m.host.cfg.dim;
        dev::upwind(dim, dm.cells.d, 0, dm.cells.d.ncells, u.d, unp1.d, a[0], a[1], dim == 3 ? a[2] : 0.0, dt, m.ctx->stream);
    }

    void upwind_scalar_burgers(const double* k, double dt, Field& u, Field& unp1)
    {
        u.check_current();
        unp1.check_current();
        if (u.mesh != unp1.mesh || &u == &unp1)
        {
            throw std::invalid_argument("upwind_scalar_burgers: u and unp1 must be two fields on one mesh");
        }
        Mesh& m = *u.mesh;
        if (m.host.cfg.dim != 2)
        {
            throw std::invalid_argument("upwind_scalar_burgers is two-dimensional (stencil_field.hpp:218-247)");
        }
        DevMesh& dm = m.device();
        ScopedTimer timer(m.ctx, "field expressions");
        dev::upwind_scalar_burgers(dm.cells.d, 0, dm.cells.d.ncells, u.d, unp1.d, k[0], k[1], dt, m.ctx->stream);
    }

    void sub_scaled(Field& u, double a, Field& v, Field& out)
    {
        u.check_current();
        v.check_current();
        out.check_current();
        if (u.mesh != v.mesh || u.mesh != out.mesh)
        {
            throw std::invalid_argument("field expression: fields on different meshes");
        }
        Mesh& m     = *u.mesh;
        DevMesh& dm = m.device();
        ScopedTimer timer(m.ctx, "field expressions");
        dev::sub_scaled(dm.cells.d, 0, dm.cells.d.ncells, u.d, a, v.d, out.d, m.ctx->stream);
    }

    void upwind_hand_loop(double dt, Field& u, Field& unp1)
    {
        u.check_current();
        unp1.check_current();
        if (u.mesh != unp1.mesh || &u == &unp1)
        {
            throw std::invalid_argument("hand loop: u and unp1 must be two fields on one mesh");
        }
        Mesh& m = *u.mesh;
        if (m.host.cfg.dim != 2)
        {
            throw std::invalid_argument("hand loop: the README loop is two-dimensional");
        }
        DevMesh& dm = m.device();
        ScopedTimer timer(m.ctx, "field expressions");
        dev::hand_loop(dm.cells.d, 0, dm.cells.d.ncells, u.d, unp1.d, dt, m.ctx->stream);
    }

    void convection_upwind(const double* v, Field& u, Field& out)
    {
        u.check_current();
        out.check_current();
        if (u.mesh != out.mesh || &u == &out)
        {
            throw std::invalid_argument("convection: input and output must be two fields on one mesh");
        }
        Mesh& m       = *u.mesh;
        const int dim = m.host.cfg.dim;
        update_ghost_mr(u);
        DevMesh& dm = m.device();
        ScopedTimer timer(m.ctx, "convection operator");
        bool same = dm.conv != nullptr;
        for (int d = 0; same && d < dim; ++d)
        {
            same = dm.conv->v[d] == v[d];
        }
        if (!same)
        {
            ConvTerms t = build_convection_terms(m.host, v);
            auto c      = std::make_unique<DevMesh::Conv>();
            for (int d = 0; d < dim; ++d)
            {
                c->v[d] = v[d];
            }
            c->blob       = std::make_unique<Blob>();
            const size_t orow = c->blob->add(t.row.data(), t.row.size() * 8);
            const size_t optr = c->blob->add(t.ptr.data(), t.ptr.size() * 8);
            const size_t osrc = c->blob->add(t.src.data(), t.src.size() * 8);
            const size_t ocof = c->blob->add(t.coef.data(), t.coef.size() * 8);
            c->blob->upload(m.ctx->stream);
            c->row   = reinterpret_cast<const int64_t*>(c->blob->dptr(orow));
            c->ptr   = reinterpret_cast<const int64_t*>(c->blob->dptr(optr));
            c->src   = reinterpret_cast<const int64_t*>(c->blob->dptr(osrc));
            c->coef  = reinterpret_cast<const double*>(c->blob->dptr(ocof));
            c->nrows = int64_t(t.row.size());
            dm.conv  = std::move(c);
        }
        hipStream_t s = m.ctx->stream;
        hip_check(hipMemsetAsync(out.d, 0, size_t(out.size) * 8, s), "hipMemsetAsync(convection)");
        dev::conv_gather(dm.conv->row, dm.conv->ptr, dm.conv->src, dm.conv->coef, dm.conv->nrows, u.d, out.d, s);
    }

    namespace
    {
        std::unique_ptr<DevMesh::Conv> upload_terms(const ConvTerms& t, hipStream_t s)
        {
            auto c            = std::make_unique<DevMesh::Conv>();
            c->blob           = std::make_unique<Blob>();
            const size_t orow = c->blob->add(t.row.data(), t.row.size() * 8);
            const size_t optr = c->blob->add(t.ptr.data(), t.ptr.size() * 8);
            const size_t osrc = c->blob->add(t.src.data(), t.src.size() * 8);
            const size_t ocof = c->blob->add(t.coef.data(), t.coef.size() * 8);
            const size_t odir = t.dir.empty() ? 0 : c->blob->add(t.dir.data(), t.dir.size());
            c->blob->upload(s);
            c->row   = reinterpret_cast<const int64_t*>(c->blob->dptr(orow));
            c->ptr   = reinterpret_cast<const int64_t*>(c->blob->dptr(optr));
            c->src   = reinterpret_cast<const int64_t*>(c->blob->dptr(osrc));
            c->coef  = reinterpret_cast<const double*>(c->blob->dptr(ocof));
            c->dir   = t.dir.empty() ? nullptr : reinterpret_cast<const int8_t*>(c->blob->dptr(odir));
            c->nrows = int64_t(t.row.size());
            return c;
        }
    } // namespace

    void convection_upwind_nonlinear(Field& u, Field& out)
    {
        u.check_current();
        out.check_current();
        if (u.mesh != out.mesh || &u == &out)
        {
            throw std::invalid_argument("convection: input and output must be two fields on one mesh");
        }
        Mesh& m = *u.mesh;
        update_ghost_mr(u);
        DevMesh& dm = m.device();
        ScopedTimer timer(m.ctx, "convection operator");
        hipStream_t s = m.ctx->stream;
        if (!dm.burgers)
        {
            dm.burgers = upload_terms(build_burgers_terms(m.host), s);
        }
        hip_check(hipMemsetAsync(out.d, 0, size_t(out.size) * 8, s), "hipMemsetAsync(convection)");
        dev::conv_burgers_gather(dm.burgers->row, dm.burgers->ptr, dm.burgers->src, dm.burgers->coef, dm.burgers->nrows, u.d, out.d, s);
    }

    void convection_weno5(const double* v, Field& u, Field& out)
    {
        u.check_current();
        out.check_current();
        if (u.mesh != out.mesh || &u == &out)
        {
            throw std::invalid_argument("convection: input and output must be two fields on one mesh");
        }
        Mesh& m       = *u.mesh;
        const int dim = m.host.cfg.dim;
        double vel[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d)
        {
            vel[d] = v[d];
        }
        DevMesh& dm   = m.device();
        hipStream_t s = m.ctx->stream;
        // the dense kernel where the storage is the dense box (every one-level
        // mesh built here), the ordered lists otherwise or when SMR_WENO5_LISTS=1
        const int dense = std::getenv("SMR_WENO5_LISTS") ? -1 : weno5_dense_level(m.host);
        if (dense < 0 && !dm.weno5)
        {
            dm.weno5 = upload_terms(build_weno5_terms(m.host), s);
        }
        update_ghost_mr(u);
        ScopedTimer timer(m.ctx, "convection operator");
        hip_check(hipMemsetAsync(out.d, 0, size_t(out.size) * 8, s), "hipMemsetAsync(convection)");
        if (dense >= 0)
        {
            const double h  = 1.0 / double(1 << dense);
            const double fa = std::pow(h, dim - 1) / std::pow(h, dim); // h_factor (flux_based_scheme__lin_hom.hpp:62-67)
            dev::weno5_dense(dim, 1 << dense, vel, fa, u.d, out.d, s);
        }
        else
        {
            const DevMesh::Conv& c = *dm.weno5;
            dev::conv_weno5_gather(c.row, c.ptr, c.src, c.coef, c.dir, c.nrows, vel, u.d, out.d, s);
        }
    }

    void convection_upwind_heterogeneous(Field* const* velocity, Field& u, Field& out)
    {
        u.check_current();
        out.check_current();
        if (u.mesh != out.mesh || &u == &out)
        {
            throw std::invalid_argument("convection: input and output must be two fields on one mesh");
        }
        Mesh& m       = *u.mesh;
        const int dim = m.host.cfg.dim;
        const double* vel[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < dim; ++d)
        {
            Field* v = velocity[d];
            if (!v || v->mesh != u.mesh || v == &out)
            {
                throw std::invalid_argument("convection: the velocity components must be fields on the mesh of u, other than the output");
            }
            v->check_current();
        }
        // FV_scheme.hpp:176-196: the parameter field's ghosts, then the input's
        for (int d = 0; d < dim; ++d)
        {
            update_ghost_mr(*velocity[d]);
            vel[d] = velocity[d]->d;
        }
        update_ghost_mr(u);
        DevMesh& dm = m.device();
        ScopedTimer timer(m.ctx, "convection operator");
        hipStream_t s = m.ctx->stream;
        if (!dm.heterogeneous)
        {
            dm.heterogeneous = upload_terms(build_heterogeneous_terms(m.host), s);
        }
        hip_check(hipMemsetAsync(out.d, 0, size_t(out.size) * 8, s), "hipMemsetAsync(convection)");
        const DevMesh::Conv& c = *dm.heterogeneous;
        dev::conv_het_gather(c.row, c.ptr, c.src, c.coef, c.dir, c.nrows, u.d, vel, out.d, s);
    }

    // ------------------------------------------------------------------
    // adaptation (mr/adapt.hpp:147-389)
    // ------------------------------------------------------------------
    namespace
    {
        std::vector<Field*> adapted_fields(AdaptState& st)
        {
            return st.fields.empty() ? std::vector<Field*>{st.field} : st.fields;
        }

        // detail arrays of all adapted fields (ncomp x n) and the tags (n)
        // the relative detail's device words (null when the detail is absolute)
        unsigned long long* relative_words(AdaptState& st)
        {
            if (!st.relative_detail)
            {
                return nullptr;
            }
            if (st.rel_capacity < st.ncomp())
            {
                if (st.rel_max)
                {
                    hip_check(hipFree(st.rel_max), "hipFree(relative detail)");
                    st.rel_max = nullptr;
                }
                hip_check(hipMalloc(reinterpret_cast<void**>(&st.rel_max), size_t(st.ncomp()) * 8), "hipMalloc(relative detail)");
                st.rel_capacity = st.ncomp();
            }
            return st.rel_max;
        }

        void ensure_adapt_buffers(AdaptState& st, int64_t n)
        {
            n *= st.ncomp();
            if (n > st.capacity)
            {
                if (st.detail)
                {
                    hip_check(hipFree(st.detail), "hipFree(detail)");
                }
                if (st.tag)
                {
                    hip_check(hipFree(st.tag), "hipFree(tag)");
                }
                st.capacity = std::max<int64_t>(n, 1);
                hip_check(hipMalloc(reinterpret_cast<void**>(&st.detail), size_t(st.capacity) * 8), "hipMalloc(detail)");
                hip_check(hipMalloc(reinterpret_cast<void**>(&st.tag), size_t(st.capacity)), "hipMalloc(tag)");
            }
        }

        // update_fields (algorithm/update_fields.hpp:27-127) into a new mesh image
        void update_fields(const std::vector<Field*>& fs, MRMeshData& nm, DevMesh& ndev)
        {
            Mesh& m         = *fs[0]->mesh;
            Ctx& ctx        = *m.ctx;
            ScopedTimer timer(&ctx, "fields update");
            hipStream_t s   = ctx.stream;
            DevMesh& odev   = *m.dev;
            DevWorkList local[3];
            if (!ndev.has_uf)
            {
                // lists built and uploaded here (the image came without them)
                build_update_fields_lists(m.host, nm, local[0].host, local[1].host, local[2].host);
                for (DevWorkList& l : local)
                {
                    upload_list(ctx, l, ndev.table, odev.table);
                }
            }
            // the resolve of the image's lists and of these is checked once
            ctx.check_resolve();
            const DevWorkList& copy    = ndev.has_uf ? ndev.uf[0] : local[0];
            const DevWorkList& coarsen = ndev.has_uf ? ndev.uf[1] : local[1];
            const DevWorkList& refine  = ndev.has_uf ? ndev.uf[2] : local[2];
            for (Field* f : fs)
            {
                double* nd = static_cast<double*>(dev_alloc(size_t(std::max<int64_t>(nm.nref, 1)) * 8, s));
                hip_check(hipMemsetAsync(nd, 0, size_t(nm.nref) * 8, s), "hipMemsetAsync(new field)");
                dev::copy_cells(copy.d, 0, copy.d.ncells, nd, f->d, s);
                dev::project(nm.cfg.dim, coarsen.d, 0, coarsen.d.ncells, nd, f->d, s);
                dev::predict_children(nm.cfg.dim, nm.cfg.pred_s, refine.d, 0, refine.d.ncells, nd, f->d, s);
                // the lists and the old storage are released in stream order after these kernels
                dev_free(f->d, s);
                f->d    = nd;
                f->size = nm.nref;
            }
        }
    } // namespace

    dev::Bits make_bits(const MeshConfig& cfg)
    {
        dev::Bits b;
        b.lo        = cfg.min_level;
        b.hi        = cfg.max_level;
        int64_t off = 0;
        for (int l = 0; l <= 20; ++l)
        {
            const int64_t n = int64_t(1) << l;
            b.wpr[l]        = std::max<int64_t>(1, n / 64);
            b.off[l]        = off;
            if (l >= b.lo && l <= b.hi)
            {
                off += b.wpr[l] * (cfg.dim == 3 ? n * n : n);
            }
        }
        b.off[21] = off;
        return b;
    }

    namespace
    {
        // the five ids of the mesh with the leaves `leaves` (unpadded bitmaps
        // of b), built on the device and read back as runs; the run buffer
        // is sized from the current mesh's runs, with a second round trip
        // only if that is too small
        // allocates the arrays of the device-built lists from their table
        // (packed runs << 32 | cells scans at the first word of every level)
        // and launches the kernel that writes them
        void emit_lists(Ctx& ctx, const dev::PBits& p, const MeshConfig& cfg, int nlists, const uint64_t* bits, const uint64_t* scan,
                        const std::vector<uint64_t>& table, DeviceLists& dl)
        {
            dl.nlists    = nlists;
            size_t bytes = 0;
            dev::ListOuts outs{};
            std::vector<size_t> off(static_cast<size_t>(nlists));
            for (int k = 0; k < nlists; ++k)
            {
                off[size_t(k)] = bytes;
                bytes += list_from_table(table.data() + size_t(k) * dev::SM_SEG, dl.l[k], outs.l[k]);
            }
            dl.mem  = std::make_shared<DevBuffer>();
            char* d = static_cast<char*>(dl.mem->get(bytes, ctx.stream));
            for (int k = 0; k < nlists; ++k)
            {
                bind_list_arrays(d + off[size_t(k)], dl.l[k], outs.l[k]);
            }
            dev::lists_emit(p, cfg.min_level, cfg.max_level, nlists, bits, scan, outs, ctx.stream);
        }

        // the five ids of the mesh with the leaves `leaves` (unpadded bitmaps
        // of b), built on the device and read back as runs; the run buffer
        // is sized from the current mesh's runs, with a second round trip
        // only if that is too small.  With dl: the work lists of the new mesh
        // too (and its update_fields lists from old_sets, the padded ids of
        // the current mesh, when given), their per-level offsets read back in
        // the same round trip; dl keeps the new padded ids.
        MRMeshData device_submesh(Ctx& ctx, const MeshConfig& cfg, const dev::Bits& b, const uint64_t* leaves, const MRMeshData& current,
                                  const uint64_t* old_sets, DeviceLists* dl)
        {
            hipStream_t s          = ctx.stream;
            const int pad          = std::max(cfg.radius, cfg.pred_s);
            const dev::PBits p     = dev::make_pbits(cfg.dim, cfg.max_level, pad);
            const int64_t total    = p.off[p.hi + 1];
            const int64_t nwords   = total * dev::SM_NSETS;
            int64_t cur_runs       = 0;
            for (const CellArray& ca : current.ids)
            {
                for (const LevelSet& ls : ca)
                {
                    cur_runs += int64_t(ls.runs.size());
                }
            }
            const int64_t cap     = cur_runs + cur_runs / 8 + 4096;
            const size_t tbytes   = (dev::submesh_scan_temp_bytes(nwords) + 255) & ~size_t(255);
            const size_t kbytes   = (size_t(std::max<int64_t>(2 * total, 1)) * 8 + 255) & ~size_t(255); // 2 scratch maps
            const size_t cbytes   = (size_t(2 * std::max<int64_t>(nwords, 1)) * 4 + 255) & ~size_t(255);
            constexpr size_t gb   = (size_t(dev::SM_NSETS * dev::SM_SEG) * 4 + 255) & ~size_t(255);
            const size_t rbytes   = size_t(dev::SM_RUN) * 4;
            DevBuffer sets_buf;
            uint64_t* sets = static_cast<uint64_t*>(sets_buf.get(size_t(std::max<int64_t>(nwords, 1)) * 8, s));
            DevBuffer work;
            char* wp         = static_cast<char*>(work.get(kbytes + cbytes + tbytes + gb + size_t(cap) * rbytes + 256, s));
            uint64_t* scr    = reinterpret_cast<uint64_t*>(wp);
            int32_t* counts  = reinterpret_cast<int32_t*>(wp + kbytes);
            int32_t* offsets = counts + std::max<int64_t>(nwords, 1);
            void* temp       = wp + kbytes + cbytes;
            int32_t* dseg    = reinterpret_cast<int32_t*>(wp + kbytes + cbytes + tbytes);
            int32_t* dr      = reinterpret_cast<int32_t*>(wp + kbytes + cbytes + tbytes + gb);
            dev::submesh_build(b, leaves, p, cfg.radius, cfg.pred_s, cfg.max_level != cfg.min_level, sets, scr, s);
            dev::submesh_run_offsets(p, dev::SM_NSETS, sets, counts, offsets, temp, tbytes, s);
            dev::submesh_run_emit(p, dev::SM_NSETS, sets, offsets, dr, cap, s);
            dev::submesh_segments(p, dev::SM_NSETS, offsets, counts, dseg, s);
            // the segment table (its last entry is the run count) and the runs
            // come back first; the work lists (bitmaps, packed counts, their
            // scan, per-level table) are built while the host completes the mesh
            const int nlists     = dl ? (old_sets ? int(dev::DL_N) : int(dev::DL_UF_COPY)) : 0;
            const int64_t lwords = nlists ? dev::lists_words(p, cfg.min_level, cfg.max_level, nlists) : 0;
            const size_t ltb     = nlists ? (dev::lists_scan_temp_bytes(lwords) + 255) & ~size_t(255) : 0;
            const size_t lgb     = (size_t(std::max(nlists, 1)) * dev::SM_SEG * 8 + 255) & ~size_t(255);
            char* h      = ctx.readback->acquire(lgb + gb + size_t(cap) * rbytes);
            uint64_t* ht = reinterpret_cast<uint64_t*>(h);
            int32_t* hs  = reinterpret_cast<int32_t*>(h + lgb);
            int32_t* hr  = reinterpret_cast<int32_t*>(h + lgb + gb);
            hip_check(hipMemcpyAsync(hs, dseg, gb, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(submesh)");
            hip_check(hipMemcpyAsync(hr, dr, size_t(cap) * rbytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(submesh)");
            ctx.mark_runs_back();
            DevBuffer lwork;
            uint64_t *lbits = nullptr, *lscan = nullptr;
            if (nlists)
            {
                // the scan's temporary storage 256-byte aligned (its look-back
                // state is read and written with wide atomics)
                const size_t lsz   = (size_t(3 * lwords) * 8 + 255) & ~size_t(255);
                char* lp           = static_cast<char*>(lwork.get(lsz + ltb + lgb + 256, s));
                lbits              = reinterpret_cast<uint64_t*>(lp);
                uint64_t* lcounts  = lbits + lwords;
                lscan              = lcounts + lwords;
                void* ltemp        = lp + lsz;
                uint64_t* ltable   = reinterpret_cast<uint64_t*>(lp + lsz + ltb);
                dev::lists_build(p, cfg.min_level, cfg.max_level, sets, old_sets, nlists, lbits, lcounts, lscan, ltemp, ltb, ltable, s);
                hip_check(hipMemcpyAsync(ht, ltable, size_t(nlists) * dev::SM_SEG * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(lists)");
            }
            hip_check(hipEventSynchronize(ctx.runs_back), "hipEventSynchronize(runs)");
            const int64_t nruns = hs[dev::SM_NSETS * dev::SM_SEG - 1];
            std::vector<uint64_t> table;
            if (nruns > cap)
            {
                // a second round trip: the list table first (the readback
                // buffer is acquired again)
                ctx.sync();
                table.assign(ht, ht + size_t(nlists) * dev::SM_SEG);
                DevBuffer rb;
                int32_t* big = static_cast<int32_t*>(rb.get(size_t(nruns) * rbytes, s));
                dev::submesh_run_emit(p, dev::SM_NSETS, sets, offsets, big, nruns, s);
                h  = ctx.readback->acquire(gb + size_t(nruns) * rbytes);
                hs = reinterpret_cast<int32_t*>(h);
                hr = reinterpret_cast<int32_t*>(h + gb);
                hip_check(hipMemcpyAsync(hs, dseg, gb, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(submesh)");
                hip_check(hipMemcpyAsync(hr, big, size_t(nruns) * rbytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(submesh)");
                ctx.sync();
            }
            MRMeshData m;
            meshes_from_runs({&cfg}, hr, hs, {&m});
            if (nlists)
            {
                // the lists' sizes and levels, then their runs
                if (table.empty())
                {
                    ctx.sync();
                    table.assign(ht, ht + size_t(nlists) * dev::SM_SEG);
                }
                emit_lists(ctx, p, cfg, nlists, lbits, lscan, table, *dl);
                dl->sets = std::move(sets_buf);
            }
            return m;
        }
    } // namespace

    bool host_lists()
    {
        static const bool v = []
        {
            const char* e = std::getenv("SMR_HOST_LISTS");
            return e && std::atoi(e) != 0;
        }();
        return v;
    }

    bool host_submesh()
    {
        static const bool v = []
        {
            const char* e = std::getenv("SMR_HOST_SUBMESH");
            return e && std::atoi(e) != 0;
        }();
        return v;
    }

    namespace
    {
        // level set `ls` from the runs [a, e): the runs are maximal and in
        // (z, y, x) order, so the rows are counted first and every array is
        // stored at its final size
        void fill_level_set(LevelSet& ls, const int32_t* runs, int64_t a, int64_t e)
        {
            const int32_t* r = runs + dev::SM_RUN * a;
            const int64_t nr = e - a;
            int64_t rows     = 1;
            for (int64_t k = 1; k < nr; ++k)
            {
                const int32_t* u = r + dev::SM_RUN * k;
                rows += (u[2] != u[2 - dev::SM_RUN]) | (u[3] != u[3 - dev::SM_RUN]);
            }
            ls.runs.resize(size_t(nr));
            ls.ry.resize(size_t(rows));
            ls.rz.resize(size_t(rows));
            ls.rp.resize(size_t(rows) + 1);
            Run* ru    = ls.runs.data();
            int32_t* y = ls.ry.data();
            int32_t* z = ls.rz.data();
            int64_t* o = ls.rp.data();
            int64_t w  = -1;
            for (int64_t k = 0; k < nr; ++k)
            {
                const int32_t* u = r + dev::SM_RUN * k;
                if (w < 0 || u[2] != y[w] || u[3] != z[w])
                {
                    ++w;
                    y[w] = u[2];
                    z[w] = u[3];
                    o[w] = k;
                }
                ru[k] = Run{u[0], u[1]};
            }
            o[rows] = nr;
        }

        void init_mesh(MRMeshData& m, const MeshConfig& cfg)
        {
            m     = MRMeshData();
            m.cfg = cfg;
            for (auto& ca : m.ids)
            {
                ca = empty_cell_array(cfg.dim);
            }
        }
    } // namespace

    void meshes_from_runs(const std::vector<const MeshConfig*>& cfgs, const int32_t* runs, const int32_t* seg, const std::vector<MRMeshData*>& out)
    {
        const int n        = int(cfgs.size());
        const int nthreads = HostPool::get().threads();
        auto segment       = [&](int q, int id, int level)
        {
            const int32_t* sg = seg + (int64_t(q) * dev::SM_NSETS + id) * dev::SM_SEG;
            return std::make_pair(int64_t(sg[level]), int64_t(sg[level + 1]));
        };
        auto levels = [&](int q) { return std::min(cfgs[size_t(q)]->max_level + 1, dev::SM_SEG - 1); };
        if (n >= 2 * nthreads)
        {
            // many meshes: one task per mesh, the largest first
            std::vector<int> order(static_cast<size_t>(n));
            for (int q = 0; q < n; ++q)
            {
                order[size_t(q)] = q;
            }
            auto size = [&](int q) { return seg[(int64_t(q) + 1) * dev::SM_NSETS * dev::SM_SEG - 1] - seg[int64_t(q) * dev::SM_NSETS * dev::SM_SEG]; };
            std::sort(order.begin(), order.end(), [&](int a, int b) { return size(a) > size(b); });
            parallel_for(n,
                         [&](int t)
                         {
                             const int q   = order[size_t(t)];
                             MRMeshData& m = *out[size_t(q)];
                             init_mesh(m, *cfgs[size_t(q)]);
                             for (int id = 0; id < dev::SM_NSETS; ++id)
                             {
                                 for (int level = 0; level < levels(q); ++level)
                                 {
                                     const auto [a, e] = segment(q, id, level);
                                     if (e > a)
                                     {
                                         fill_level_set(m.ids[size_t(id)][size_t(level)], runs, a, e);
                                     }
                                 }
                             }
                             complete_mesh(m);
                         });
            return;
        }
        // few meshes: every non-empty level set is a segment of the runs; the
        // segments are split in chunks of about equal run counts, one task each
        // (a task per segment would contend on the pool's task counter)
        parallel_for(n, [&](int q) { init_mesh(*out[size_t(q)], *cfgs[size_t(q)]); });
        struct Seg
        {
            int q, id, level;
            int64_t a, e;
        };
        std::vector<Seg> segs;
        std::vector<int64_t> work{0};
        for (int q = 0; q < n; ++q)
        {
            for (int id = 0; id < dev::SM_NSETS; ++id)
            {
                for (int level = 0; level < levels(q); ++level)
                {
                    const auto [a, e] = segment(q, id, level);
                    if (e > a)
                    {
                        segs.push_back({q, id, level, a, e});
                        work.push_back(work.back() + (e - a) + 16);
                    }
                }
            }
        }
        const int nchunks = int(std::min<size_t>(segs.size(), size_t(8 * nthreads)));
        parallel_for(nchunks,
                     [&](int c)
                     {
                         // the segments whose work prefix starts in [lo, hi)
                         const int64_t lo = work.back() * c / nchunks, hi = work.back() * (c + 1) / nchunks;
                         const size_t s0  = size_t(std::lower_bound(work.begin(), work.end() - 1, lo) - work.begin());
                         const size_t s1  = size_t(std::lower_bound(work.begin(), work.end() - 1, hi) - work.begin());
                         for (size_t t = s0; t < s1; ++t)
                         {
                             const Seg& g = segs[t];
                             fill_level_set(out[size_t(g.q)]->ids[size_t(g.id)][size_t(g.level)], runs, g.a, g.e);
                         }
                     });
        parallel_for(n, [&](int q) { complete_mesh(*out[size_t(q)]); });
    }

    bool device_mesh_update(Ctx& ctx, DevMesh& dm, const MeshConfig& cfg, const uint8_t* tag, CellArray& out, const MRMeshData* current,
                            MRMeshData* mesh, DeviceLists* lists)
    {
        hipStream_t s       = ctx.stream;
        const dev::Bits b   = make_bits(cfg);
        const int64_t nw    = b.off[b.hi + 1];
        const size_t wbytes = size_t(std::max<int64_t>(nw, 1)) * 8;
        if (!dm.has_leaf_bits)
        {
            dev::bits_from_cells(dm.cells.d, b, static_cast<uint64_t*>(dm.leaf_bits.get(wbytes, s)), s);
            dm.has_leaf_bits = true;
        }
        const uint64_t* old = static_cast<const uint64_t*>(dm.leaf_bits.p);
        // graduation passes launched per readback of their flags: the fixed
        // point is the first pass that refines nothing, which on the 2D
        // L4..12 advection was the first or the second pass of every update
        // (a later one costs another readback)
        constexpr int kPasses = 2;
        DevBuffer work;
        uint64_t* w   = static_cast<uint64_t*>(work.get(5 * wbytes + 256, s));
        uint64_t* add = w;
        uint64_t* rem = w + nw;
        uint64_t* buf[2] = {w + 2 * nw, w + 3 * nw};
        uint64_t* q   = w + 4 * nw;
        int* flags    = reinterpret_cast<int*>(w + 5 * nw); // kPasses pass flags, then the differ flag
        int h_flags[kPasses + 1];
        {
            ScopedTimer t(&ctx, "update_cell_array_from_tag");
            dev::bits_tag_marks(cfg.dim, dm.cells.d, b, tag, cfg.min_level, cfg.max_level, add, rem, s);
            dev::bits_combine(b, old, add, rem, buf[0], s);
        }
        int cur = 0;
        {
            ScopedTimer t(&ctx, "make_graduation");
            // make_graduation's fixed point (graduation.hpp:623-723): passes
            // until one refines nothing.  A pass that refines nothing copies
            // its input, so the passes after it leave the leaves as they are
            // and the last buffer holds the fixed point; the difference with
            // the old leaves is taken in the same round trip.
            for (int batch = 0;; ++batch)
            {
                if (batch > 21)
                {
                    throw std::logic_error("make_graduation: no fixed point");
                }
                hip_check(hipMemsetAsync(flags, 0, sizeof(int) * (kPasses + 1), s), "hipMemsetAsync(flags)");
                for (int p = 0; p < kPasses; ++p)
                {
                    dev::bits_graduation_pass(cfg.dim, b, buf[cur], q, buf[1 - cur], 2 * cfg.grad_width, flags + p, s);
                    cur = 1 - cur;
                }
                dev::bits_differ(b, buf[cur], old, flags + kPasses, s);
                hip_check(hipMemcpyAsync(h_flags, flags, sizeof(h_flags), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(flags)");
                ctx.sync();
                bool fixed = false;
                for (int p = 0; p < kPasses; ++p)
                {
                    fixed = fixed || h_flags[p] == 0;
                }
                if (fixed)
                {
                    break;
                }
            }
        }
        if (!h_flags[kPasses])
        {
            return false;
        }
        if (mesh && current && !host_submesh())
        {
            ScopedTimer t(&ctx, "mesh construction");
            DevBuffer old_buf;
            const uint64_t* old_sets = nullptr;
            if (lists && dm.has_mesh_sets)
            {
                old_sets = static_cast<const uint64_t*>(dm.mesh_sets.p);
            }
            else if (lists)
            {
                // the padded ids of the current mesh, from its leaves
                const dev::PBits p  = dev::make_pbits(cfg.dim, cfg.max_level, std::max(cfg.radius, cfg.pred_s));
                const int64_t total = p.off[p.hi + 1];
                uint64_t* o = static_cast<uint64_t*>(old_buf.get(size_t(std::max<int64_t>((dev::SM_NSETS + 2) * total, 1)) * 8, s));
                dev::submesh_build(b, old, p, cfg.radius, cfg.pred_s, cfg.max_level != cfg.min_level, o, o + dev::SM_NSETS * total, s);
                old_sets = o;
            }
            *mesh           = device_submesh(ctx, cfg, b, buf[cur], *current, old_sets, lists);
            mesh->graduated = true;
            return true;
        }
        ScopedTimer t(&ctx, "leaves to host");
        // run-length encoding of the new leaves, in traversal order, into a
        // buffer sized from the old leaves' runs; one round trip brings the run
        // count and the runs (a second one only if the buffer was too small)
        const uint64_t* leaves = buf[cur];
        const int64_t nl       = nw - b.off[b.lo];
        const size_t tbytes    = dev::bits_scan_temp_bytes(b);
        const int64_t cap      = dm.cells.d.nruns + dm.cells.d.nruns / 4 + 1024;
        const size_t cbytes    = (size_t(2 * std::max<int64_t>(nl, 1)) * 4 + 255) & ~size_t(255);
        DevBuffer scan;
        char* sp          = static_cast<char*>(scan.get(cbytes + tbytes + size_t(cap) * 20, s));
        int32_t* counts   = reinterpret_cast<int32_t*>(sp);
        int32_t* offsets  = counts + std::max<int64_t>(nl, 1);
        void* temp        = sp + cbytes;
        int32_t* dr       = reinterpret_cast<int32_t*>(sp + cbytes + tbytes);
        dev::bits_run_offsets(cfg.dim, b, leaves, counts, offsets, temp, tbytes, s);
        dev::bits_run_emit(cfg.dim, b, leaves, offsets, dr, cap, s);
        int32_t last[2] = {0, 0};
        std::vector<int32_t> runs(size_t(cap) * 5);
        hip_check(hipMemcpyAsync(&last[0], offsets + nl - 1, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(runs)");
        hip_check(hipMemcpyAsync(&last[1], counts + nl - 1, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(runs)");
        hip_check(hipMemcpyAsync(runs.data(), dr, size_t(cap) * 20, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(runs)");
        ctx.sync();
        const int64_t nruns = int64_t(last[0]) + last[1];
        if (nruns > cap)
        {
            DevBuffer rb;
            int32_t* big = static_cast<int32_t*>(rb.get(size_t(nruns) * 20, s));
            dev::bits_run_emit(cfg.dim, b, leaves, offsets, big, nruns, s);
            runs.resize(size_t(nruns) * 5);
            hip_check(hipMemcpyAsync(runs.data(), big, size_t(nruns) * 20, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(runs)");
            ctx.sync();
        }
        out = empty_cell_array(cfg.dim);
        for (int64_t k = 0; k < nruns; ++k)
        {
            const int32_t* r = runs.data() + 5 * k;
            out[size_t(r[4])].push(r[2], r[3], r[0], r[1]);
        }
        return true;
    }

    dev::Thresholds make_thresholds(const MeshConfig& cfg, int ite, double eps, double regularity)
    {
        const int dim = cfg.dim, minl = cfg.min_level, maxl = cfg.max_level;
        dev::Thresholds thr{};
        for (int level = minl; level <= maxl - ite; ++level)
        {
            const int exponent = dim * (maxl - level);
            const double eps_l = eps / (1 << exponent);
            const double reg   = regularity + dim;
            const double fine  = std::pow(2.0, reg) * eps_l;
            thr.eps[level]     = eps_l;
            thr.fine[level]    = fine;
            thr.coarse[level]  = fine / (1 << dim);
        }
        return thr;
    }

    void harten_tags(Field& f, int ite, double eps, double regularity, uint8_t* tag, double* det)
    {
        Field* fs[1] = {&f};
        harten_tags(fs, 1, ite, eps, regularity, tag, det);
    }

    void harten_tags(Field* const* fs, int nf, int ite, double eps, double regularity, uint8_t* tag, double* det,
                     unsigned long long* rel_max, bool refine_boundary)
    {
        Field& f = *fs[0];
        for (int k = 0; k < nf; ++k)
        {
            Field* g = fs[k];
            g->check_current();
            if (g->mesh != f.mesh)
            {
                throw std::invalid_argument("make_MRAdapt: the fields must live on one mesh");
            }
        }
        Mesh& m              = *f.mesh;
        Ctx& ctx             = *m.ctx;
        hipStream_t s        = ctx.stream;
        const MeshConfig cfg = m.host.cfg;
        const int dim = cfg.dim, minl = cfg.min_level, maxl = cfg.max_level;
        DevMesh& dm     = m.device_adapt();
        const int64_t n   = m.host.nref;
        const int ncomp   = nf;
        dev::zero_f64(det, n * ncomp, s);
        dev::zero_u8(tag, n, s);
        dev::tag_init(dm.cells.d, 0, dm.cells.d.ncells, tag, s);

        // update_ghost_mr(m_fields) (mr/adapt.hpp:292)
        for (int k = 0; k < nf; ++k)
        {
            update_ghost_mr(*fs[k]);
        }

        // detail computation: levels [min-1, max-ite), one detail array per field
        {
            ScopedTimer timer(&ctx, "detail computation");
            int64_t c0, c1;
            dm.detail.cells_of_levels(minl > 0 ? minl - 1 : 0, maxl - ite, c0, c1);
            for (int k = 0; k < ncomp; ++k)
            {
                dev::detail(dim, cfg.pred_s, dm.detail.d, c0, c1, fs[size_t(k)]->d, det + k * n, s);
            }
            // compute_relative_detail (mr/adapt.hpp:319-322)
            if (rel_max)
            {
                for (int k = 0; k < ncomp; ++k)
                {
                    dev::relative_detail(dm.cells.d, 0, dm.cells.d.ncells, fs[size_t(k)]->d, det + k * n, n, rel_max + k, s);
                }
            }
        }
        // tag cells: fine levels [min, max-ite], i.e. coarse levels [min-1, max-ite-1]
        {
            ScopedTimer timer(&ctx, "tag cells");
            const dev::Thresholds thr = make_thresholds(cfg, ite, eps, regularity);
            int64_t c0, c1;
            dm.criteria.cells_of_levels(minl - 1, maxl - ite, c0, c1);
            dev::criteria(dim, dm.criteria.d, c0, c1, det, tag, thr, minl, maxl, s, ncomp, n);
        }
        // --refine-boundary (mr/adapt.hpp:341-346): the leaves of max_level
        // within the stencil radius of a face are kept
        if (refine_boundary)
        {
            int64_t c0, c1;
            dm.cells.cells_of_levels(maxl, maxl + 1, c0, c1);
            dev::keep_boundary(dm.cells.d, c0, c1, tag, dim, 1 << maxl, cfg.radius, s);
        }
        // tag finalization: levels max..1, sequential
        {
            ScopedTimer timer(&ctx, "tag finalization");
            const SeqPlan& plan = max_plan(ctx, dm, cfg);
            const auto* d_items = static_cast<const dev::SeqItem*>(plan.d_items.p);
            for (const SeqPlan::Step& st : plan.steps)
            {
                if (st.seq)
                {
                    dev::sequence(dim, cfg.pred_s, d_items, st.i0, st.i1, dev::List{}, dev::List{}, dm.maximum.d, nullptr, tag, false, 0.0, s);
                }
                else
                {
                    const dev::SeqItem& it = plan.items[size_t(st.i0)];
                    dev::maximum(dim, dm.maximum.d, it.c0, it.c1, tag, s);
                }
            }
        }
    }

    void adapt_tags(AdaptState& st, int ite, double eps, double regularity, uint8_t* tag)
    {
        Field& f = *st.field;
        f.check_current();
        ensure_adapt_buffers(st, f.mesh->host.nref);
        const std::vector<Field*> fs = adapted_fields(st);
        harten_tags(fs.data(), int(fs.size()), ite, eps, regularity, tag, st.detail, relative_words(st), st.refine_boundary);
        st.last_size      = 0; // the tags are the caller's, not the adaptation's
        st.host_tag_valid = false;
    }

    bool harten(AdaptState& st, int ite, double eps, double regularity)
    {
        Field& f        = *st.field;
        f.check_current();
        Mesh& m         = *f.mesh;
        Ctx& ctx        = *m.ctx;
        hipStream_t s   = ctx.stream;
        const MeshConfig cfg = m.host.cfg;
        const int minl = cfg.min_level, maxl = cfg.max_level;
        const int64_t n = m.host.nref;
        ensure_adapt_buffers(st, n);
        const std::vector<Field*> fs = adapted_fields(st);
        harten_tags(fs.data(), int(fs.size()), ite, eps, regularity, st.tag, st.detail, relative_words(st), st.refine_boundary);
        DevMesh& dm = *m.dev;
        std::unique_ptr<DevMesh> ndev;
        MRMeshData nm;
        DeviceLists dl;
        {
            ScopedTimer timer(&ctx, "mesh update");
            st.last_size      = n;
            st.host_tag_valid = false;
            // mesh update (graduation.hpp:743-843, 573-726): on the device,
            // the host receives the new leaves only when they change (one
            // round trip when they do not); the new mesh's ids and work lists
            // are built on the device too
            CellArray new_ca;
            nm.nref = -1; // set when the device builds the whole mesh
            if (!device_mesh_update(ctx, dm, cfg, st.tag, new_ca, &m.host, &nm, host_lists() ? nullptr : &dl))
            {
                return true;
            }
            if (nm.nref < 0)
            {
                ScopedTimer t3(nullptr, "mesh construction");
                nm           = build_mesh(cfg, std::move(new_ca));
                nm.graduated = true;
            }
        }
        if (dl.nlists == int(dev::DL_N))
        {
            // the outer-ghost recipes and the cells list on the host, then one
            // allocation for the image; the resolve is checked with the
            // update_fields lists
            std::vector<Recipes> outer;
            WorkList cells;
            {
                ScopedTimer t4(nullptr, "program build");
                build_recipes_and_cells(nm, outer, cells);
            }
            ScopedTimer timer(&ctx, "device upload");
            ndev = upload_image_device_lists(ctx, nm, std::move(outer), std::move(cells), std::move(dl), &dm.table);
        }
        else
        {
            // the complete image of the new mesh (cells, ghost and adapt
            // programs) in one allocation; its resolve is checked together with
            // the update_fields lists
            HostImage img;
            {
                // the programs and the update_fields lists from the current mesh
                ScopedTimer t4(nullptr, "program build");
                img = build_host_image(nm, &m.host);
            }
            ScopedTimer timer(&ctx, "device upload");
            ndev = upload_image(ctx, nm, std::move(img), false, &dm.table);
        }
        update_fields(fs, nm, *ndev);
        m.replace(std::move(nm), std::move(ndev));
        for (Field* g : fs)
        {
            g->version = m.version;
        }
        return false;
    }

    int adapt(AdaptState& st, double eps, double regularity, int max_iterations)
    {
        Mesh& m        = *st.field->mesh;
        const int minl = m.host.cfg.min_level, maxl = m.host.cfg.max_level;
        if (minl == maxl)
        {
            return 0;
        }
        ScopedTimer timer(m.ctx, "mesh adaptation");
        int limit = maxl - minl;
        if (max_iterations >= 0)
        {
            limit = std::min(limit, max_iterations);
        }
        for (int i = 0; i < limit; ++i)
        {
            if (harten(st, i, eps, regularity))
            {
                return i + 1;
            }
        }
        return limit;
    }
} // namespace smr
