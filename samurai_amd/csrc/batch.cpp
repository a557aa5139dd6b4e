#include "batch.hpp"

#include "host/parallel.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <optional>
#include <sstream>
#include <stdexcept>

namespace smr
{
    namespace
    {
        constexpr int AUX_THR  = 1;
        constexpr int AUX_TABS = 2;

        size_t align_up(size_t n, size_t a)
        {
            return (n + a - 1) / a * a;
        }

        // MODE_BLOCKS: a stage where some chain has more cells than this runs
        // grid-wide; the others run one block per chain (SMR_BLOCK_STAGE_MAX)
        int64_t block_stage_max()
        {
            static const int64_t v = []
            {
                const char* e = std::getenv("SMR_BLOCK_STAGE_MAX");
                const long long x = e ? std::atoll(e) : 0;
                return x > 0 ? int64_t(x) : int64_t(4096);
            }();
            return v;
        }
        constexpr int kBlockThreads = 1024;
    } // namespace

    dev::Segment make_seg(int op, const MeshConfig& cfg)
    {
        dev::Segment g;
        std::memset(&g, 0, sizeof(g));
        g.op        = op;
        g.dim       = cfg.dim;
        g.pred_s    = cfg.pred_s;
        g.min_level = cfg.min_level;
        g.max_level = cfg.max_level;
        return g;
    }

    dev::Segment list_seg(int op, const MeshConfig& cfg, const DevWorkList& l, int64_t c0, int64_t c1)
    {
        dev::Segment g = make_seg(op, cfg);
        g.list         = l.d;
        g.c0           = c0;
        g.c1           = c1;
        return g;
    }

    void push_stage(Chain& ch, const dev::Segment& g)
    {
        if (g.c1 > g.c0)
        {
            ch.push_back(Stage{g});
        }
    }

    // update_ghost_mr (update_ghost_mr.hpp:194-237) on one instance, in
    // the order of the standalone launches: per level from the finest, the
    // outer-ghost recipe stages then the projection to the level below;
    // then the prediction levels from the coarsest
    void ghost_chain(Chain& ch, DevMesh& dm, const MeshConfig& cfg, double* u, bool has_bc, double bc)
    {
        for (int level = cfg.max_level; level >= 0; --level)
        {
            DevRecipes& r  = dm.outer[size_t(level)];
            const auto& sp = r.host.stage_ptr;
            if (!sp.empty() && size_t(sp.back()) > r.host.size())
            {
                throw std::logic_error("recipe stages exceed the recipe list");
            }
            for (size_t k = 0; k + 1 < sp.size(); ++k)
            {
                if (has_bc && r.host.stage_without_bc_only(k))
                {
                    continue; // copies for B.C.-less fields only
                }
                dev::Segment g = make_seg(dev::SEG_RECIPE, cfg);
                g.rdst         = r.dst;
                g.rop          = r.op;
                g.rsp          = r.sp;
                g.rsrc         = r.src;
                g.c0           = sp[k];
                g.c1           = sp[k + 1];
                g.u            = u;
                g.has_bc       = has_bc ? 1 : 0;
                g.bc           = bc;
                push_stage(ch, g);
            }
            if (level > 0)
            {
                int64_t c0, c1;
                dm.projection.cells_of_levels(level - 1, level, c0, c1);
                dev::Segment g = list_seg(dev::SEG_PROJECT, cfg, dm.projection, c0, c1);
                g.u            = u;
                push_stage(ch, g);
            }
        }
        for (int level = 1; level <= cfg.max_level; ++level)
        {
            int64_t c0, c1;
            dm.prediction.cells_of_levels(level, level + 1, c0, c1);
            dev::Segment g = list_seg(dev::SEG_PREDICT_FINE, cfg, dm.prediction, c0, c1);
            g.u            = u;
            push_stage(ch, g);
        }
    }

    // stage k of the result is stage k of a and of b
    Chain zip(Chain a, const Chain& b)
    {
        if (a.size() < b.size())
        {
            a.resize(b.size());
        }
        for (size_t k = 0; k < b.size(); ++k)
        {
            a[k].insert(a[k].end(), b[k].begin(), b[k].end());
        }
        return a;
    }

    namespace
    {
        std::string gpu_arch(int device)
        {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, device) != hipSuccess)
            {
                return "unknown";
            }
            std::string a = prop.gcnArchName;
            const size_t c = a.find(':');
            return c == std::string::npos ? a : a.substr(0, c);
        }

        int cu_count()
        {
            static int n = -1;
            if (n < 0)
            {
                int dev = 0;
                (void)hipGetDevice(&dev);
                if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
                {
                    n = 1;
                }
            }
            return n;
        }

        int max_blocks_cached(int block)
        {
            static std::map<int, int> cache;
            auto it = cache.find(block);
            if (it == cache.end())
            {
                it = cache.emplace(block, dev::program_max_blocks(block)).first;
            }
            return it->second;
        }

        // host-side bounds of a segment before it reaches the device
        void check_segment(const dev::Segment& g)
        {
            bool ok = g.c0 >= 0 && g.c0 < g.c1;
            switch (g.op)
            {
                case dev::SEG_RECIPE: ok = ok && g.rdst && g.rop && g.rsp && g.rsrc && g.u; break;
                case dev::SEG_RESOLVE:
                    ok = ok && g.list.base && g.c1 <= dev::resolve_range(g.list) && (g.tabs || (g.aux & AUX_TABS)) && g.err;
                    break;
                case dev::SEG_FILL_F64: ok = ok && g.out; break;
                case dev::SEG_FILL_U8: ok = ok && g.tag; break;
                case dev::SEG_TAG_INIT:
                case dev::SEG_MAXIMUM: ok = ok && g.list.iv && g.c1 <= g.list.ncells && g.tag; break;
                case dev::SEG_TAG_CHANGE: ok = ok && g.list.iv && g.c1 <= g.list.ncells && g.tag && g.err; break;
                case dev::SEG_WALK_ROWS: ok = ok && g.list.walks && g.list.walk_rows && g.c1 <= g.list.nwalks; break;
                case dev::SEG_CRITERIA:
                    ok = ok && g.list.iv && g.c1 <= g.list.ncells && g.tag && g.out && (g.thr || (g.aux & AUX_THR));
                    break;
                case dev::SEG_PROJECT:
                case dev::SEG_PREDICT_FINE: ok = ok && g.list.iv && g.c1 <= g.list.ncells && g.u; break;
                case dev::SEG_DETAIL:
                case dev::SEG_UPWIND:
                case dev::SEG_COPY:
                case dev::SEG_PROJECT_TO:
                case dev::SEG_PREDICT_CHILDREN:
                case dev::SEG_CELLS_TO_COMPACT: ok = ok && g.list.iv && g.c1 <= g.list.ncells && g.u && g.out; break;
                case dev::SEG_CELLS_FROM_COMPACT: ok = ok && g.list.iv && g.c1 <= g.list.ncells && g.in2 && g.out; break;
                case dev::SEG_COPY_RAW: ok = ok && g.u && g.out; break;
                case dev::SEG_ADJ_GATHER: ok = ok && g.rdst && g.rsp && g.rsrc && g.thr && g.out; break;
                case dev::SEG_ADJ_ZERO: ok = ok && g.rdst && g.out; break;
                case dev::SEG_ADJ_UPWIND: ok = ok && g.list.iv && g.c1 <= g.list.ncells && g.u && g.out && g.in2 && g.out2; break;
                case dev::SEG_ADJ_SCATTER:
                    ok = ok && g.list.iv && g.c1 <= g.list.ncells && g.u && g.out && g.pad >= 0 && g.pad < 2 * g.dim;
                    break;
                default: ok = false;
            }
            ok = ok && (g.dim == 2 || g.dim == 3) && (g.pred_s == 0 || g.pred_s == 1);
            if (!ok)
            {
                throw std::logic_error("stage program segment out of bounds (op " + std::to_string(g.op) + ")");
            }
        }

        struct Rebuild
        {
            bool converged = false;
            MRMeshData nm;
            HostImage img;
            WorkList copy, coarsen, refine;
            // work lists built on the device: the host adds the recipes and
            // the cells list
            bool device_lists = false;
            std::array<DeviceLists::One, dev::DL_N> dl;
            std::vector<Recipes> outer;
            WorkList cells;
        };

        bool batch_host_rebuild()
        {
            static const bool v = []
            {
                const char* e = std::getenv("SMR_BATCH_HOST_REBUILD");
                return e && std::atoi(e) != 0;
            }();
            return v;
        }

        // the arrays of the device-built lists of the meshes sel[q] from their
        // table (packed runs << 32 | cells scans at the first word of every
        // level of every list), one allocation for all, and the launch that
        // writes them
        void emit_batch_lists(Ctx& ctx, const dev::SubmeshItem* d_items, const int64_t* d_cp, int64_t nwords, int n, const uint64_t* bits,
                              const uint64_t* scan, const std::vector<uint64_t>& table, const std::vector<int>& sel,
                              std::vector<std::array<DeviceLists::One, dev::DL_N>>& lists, std::shared_ptr<DevBuffer>& mem)
        {
            hipStream_t s = ctx.stream;
            std::vector<dev::ListOut> outs(size_t(n) * dev::DL_N);
            std::vector<size_t> off(size_t(n) * dev::DL_N);
            size_t bytes = 0;
            for (int q = 0; q < n; ++q)
            {
                std::array<DeviceLists::One, dev::DL_N>& ls = lists[size_t(sel[size_t(q)])];
                for (int k = 0; k < dev::DL_N; ++k)
                {
                    const size_t e = size_t(q) * dev::DL_N + size_t(k);
                    off[e]         = bytes;
                    bytes += list_from_table(table.data() + e * dev::SM_SEG, ls[size_t(k)], outs[e]);
                }
            }
            mem     = std::make_shared<DevBuffer>();
            char* d = static_cast<char*>(mem->get(bytes, s));
            for (int q = 0; q < n; ++q)
            {
                std::array<DeviceLists::One, dev::DL_N>& ls = lists[size_t(sel[size_t(q)])];
                for (int k = 0; k < dev::DL_N; ++k)
                {
                    const size_t e = size_t(q) * dev::DL_N + size_t(k);
                    bind_list_arrays(d + off[e], ls[size_t(k)], outs[e]);
                }
            }
            // the destinations to the device, then the runs
            const size_t ob = outs.size() * sizeof(dev::ListOut);
            char* h         = ctx.staging->acquire(ob);
            std::memcpy(h, outs.data(), ob);
            DevBuffer dout;
            auto* d_outs = static_cast<dev::ListOut*>(dout.get(ob, s));
            hip_check(hipMemcpyAsync(d_outs, h, ob, hipMemcpyHostToDevice, s), "hipMemcpyAsync(list destinations)");
            ctx.staging->release(s);
            dev::lists_batch_emit(d_items, d_cp, nwords, n, bits, scan, d_outs, s);
        }

        // update_sub_mesh_impl (mr/mesh.hpp:221-455) of the meshes ms[j] with
        // changed[j], from their new leaves (the concatenated unpadded bitmaps
        // laid out by items), all meshes in each launch (submesh_batch.hip);
        // the five ids come back as runs in one round trip (a second one only
        // if the buffer, sized from the current meshes' runs, is too small),
        // and meshes[j] is the complete new mesh
        //
        // With lists: the work lists of those meshes too (lists.hip, all
        // meshes per launch; the update_fields lists from the current meshes,
        // whose padded ids are built from old_leaves), (*lists)[j] those of
        // mesh j, their arrays in *list_mem.
        void construct_batch(Ctx& ctx, const std::vector<const Mesh*>& ms, const std::vector<dev::MeshItem>& items, const std::vector<bool>& changed,
                             const uint64_t* leaves, std::vector<MRMeshData>& meshes, const uint64_t* old_leaves = nullptr,
                             std::vector<std::array<DeviceLists::One, dev::DL_N>>* lists = nullptr, std::shared_ptr<DevBuffer>* list_mem = nullptr)
        {
            ScopedTimer t(&ctx, "mesh construction");
            hipStream_t s = ctx.stream;
            std::vector<int> sel;
            for (size_t j = 0; j < changed.size(); ++j)
            {
                if (changed[j])
                {
                    sel.push_back(int(j));
                }
            }
            const int n = int(sel.size());
            if (n == 0)
            {
                return;
            }
            std::vector<dev::SubmeshItem> si(static_cast<size_t>(n));
            std::vector<int64_t> pp(size_t(n) + 1, 0);
            int64_t cap = 4096;
            for (int q = 0; q < n; ++q)
            {
                const size_t j        = size_t(sel[size_t(q)]);
                const MeshConfig& c   = ms[j]->host.cfg;
                dev::SubmeshItem& it  = si[size_t(q)];
                it.b                  = items[j].b;
                it.lw0                = items[j].w0;
                it.p                  = dev::make_pbits(c.dim, c.max_level, std::max(c.radius, c.pred_s));
                it.pw0                = pp[size_t(q)];
                it.r                  = c.radius;
                it.s                  = c.pred_s;
                it.mr                 = c.max_level != c.min_level ? 1 : 0;
                it.minl               = c.min_level;
                pp[size_t(q) + 1]     = pp[size_t(q)] + it.p.off[it.p.hi + 1];
                int64_t cur           = 0;
                for (const CellArray& ca : ms[j]->host.ids)
                {
                    for (const LevelSet& ls : ca)
                    {
                        cur += int64_t(ls.runs.size());
                    }
                }
                cap += cur + cur / 8;
            }
            const int64_t PW = pp[size_t(n)];
            const int64_t nw = dev::SM_NSETS * PW;
            // the work lists' words of every item (each list over its own levels)
            std::vector<int64_t> cp(size_t(n) + 1, 0);
            for (int q = 0; q < n; ++q)
            {
                const dev::SubmeshItem& it = si[size_t(q)];
                cp[size_t(q) + 1]          = cp[size_t(q)] + (lists ? dev::lists_words(it.p, it.minl, it.p.hi, dev::DL_N) : 0);
            }
            std::optional<ScopedTimer> td(std::in_place, &ctx, "mesh ids on device");
            // the item table and its prefixes, one upload
            const size_t ib = (size_t(n) * sizeof(dev::SubmeshItem) + 255) & ~size_t(255);
            const size_t pb = size_t(n + 1) * 8;
            char* h         = ctx.staging->acquire(ib + 2 * pb);
            std::memcpy(h, si.data(), size_t(n) * sizeof(dev::SubmeshItem));
            std::memcpy(h + ib, pp.data(), pb);
            std::memcpy(h + ib + pb, cp.data(), pb);
            DevBuffer tab;
            char* dt = static_cast<char*>(tab.get(ib + 2 * pb, s));
            hip_check(hipMemcpyAsync(dt, h, ib + 2 * pb, hipMemcpyHostToDevice, s), "hipMemcpyAsync(submesh items)");
            ctx.staging->release(s);
            const auto* d_items = reinterpret_cast<const dev::SubmeshItem*>(dt);
            const auto* d_pp    = reinterpret_cast<const int64_t*>(dt + ib);
            const auto* d_cp    = reinterpret_cast<const int64_t*>(dt + ib + pb);
            // the sets, two scratch maps per mesh, run counts and offsets, scan
            // temp, the segment table, runs
            const size_t tbytes = (dev::submesh_scan_temp_bytes(nw) + 255) & ~size_t(255);
            const size_t sbytes = (size_t(nw + 2 * PW) * 8 + 255) & ~size_t(255);
            const size_t cbytes = (size_t(2 * nw) * 4 + 255) & ~size_t(255);
            const size_t gb     = (size_t(n) * dev::SM_NSETS * dev::SM_SEG * 4 + 255) & ~size_t(255);
            const size_t rbytes = size_t(dev::SM_RUN) * 4;
            DevBuffer work;
            char* wp         = static_cast<char*>(work.get(sbytes + cbytes + tbytes + gb + size_t(cap) * rbytes + 256, s));
            uint64_t* sets   = reinterpret_cast<uint64_t*>(wp);
            int32_t* counts  = reinterpret_cast<int32_t*>(wp + sbytes);
            int32_t* offsets = counts + nw;
            void* temp       = wp + sbytes + cbytes;
            int32_t* dseg    = reinterpret_cast<int32_t*>(wp + sbytes + cbytes + tbytes);
            int32_t* dr      = reinterpret_cast<int32_t*>(wp + sbytes + cbytes + tbytes + gb);
            dev::submesh_batch_build(d_items, d_pp, PW, n, leaves, sets, sets + nw, s);
            dev::submesh_batch_run_offsets(d_items, d_pp, PW, n, sets, counts, offsets, temp, tbytes, s);
            dev::submesh_batch_run_emit(d_items, d_pp, PW, n, sets, offsets, dr, cap, s);
            dev::submesh_batch_segments(d_items, d_pp, n, offsets, counts, dseg, s);
            // the segment table (its last entry is the run count) and the runs
            // come back first; the work lists (the current meshes' padded ids
            // from their leaves, then the lists' bitmaps, packed counts, scan
            // and per-level table) are built while the host completes the meshes
            const bool with_lists = lists && old_leaves;
            const int64_t lw      = with_lists ? cp[size_t(n)] : 0;
            const size_t ltb      = with_lists ? (dev::lists_scan_temp_bytes(lw) + 255) & ~size_t(255) : 0;
            const size_t lgb      = with_lists ? (size_t(n) * dev::DL_N * dev::SM_SEG * 8 + 255) & ~size_t(255) : 0;
            char* hr      = ctx.readback->acquire(lgb + gb + size_t(cap) * rbytes);
            uint64_t* ht  = reinterpret_cast<uint64_t*>(hr);
            int32_t* hs   = reinterpret_cast<int32_t*>(hr + lgb);
            int32_t* runs = reinterpret_cast<int32_t*>(hr + lgb + gb);
            hip_check(hipMemcpyAsync(hs, dseg, gb, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(submesh runs)");
            hip_check(hipMemcpyAsync(runs, dr, size_t(cap) * rbytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(submesh runs)");
            ctx.mark_runs_back();
            DevBuffer lwork;
            uint64_t *lbits = nullptr, *lscan = nullptr;
            if (with_lists)
            {
                // the scan's temporary storage 256-byte aligned (its look-back
                // state is read and written with wide atomics)
                const size_t lsz  = (size_t(nw + 2 * PW + 3 * lw) * 8 + 255) & ~size_t(255);
                char* lp          = static_cast<char*>(lwork.get(lsz + ltb + lgb + 256, s));
                uint64_t* osets   = reinterpret_cast<uint64_t*>(lp);
                lbits             = osets + nw + 2 * PW;
                uint64_t* lcounts = lbits + lw;
                lscan             = lcounts + lw;
                void* ltemp       = lp + lsz;
                uint64_t* ltable  = reinterpret_cast<uint64_t*>(lp + lsz + ltb);
                dev::submesh_batch_build(d_items, d_pp, PW, n, old_leaves, osets, osets + nw, s);
                dev::lists_batch_build(d_items, d_cp, lw, n, sets, osets, lbits, lcounts, lscan, ltemp, ltb, ltable, s);
                hip_check(hipMemcpyAsync(ht, ltable, size_t(n) * dev::DL_N * dev::SM_SEG * 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(lists)");
            }
            hip_check(hipEventSynchronize(ctx.runs_back), "hipEventSynchronize(runs)");
            const int64_t nruns = hs[int64_t(n) * dev::SM_NSETS * dev::SM_SEG - 1];
            std::vector<uint64_t> table;
            if (nruns > cap)
            {
                // a second round trip: the list table first (the readback
                // buffer is acquired again)
                ctx.sync();
                table.assign(ht, ht + (with_lists ? size_t(n) * dev::DL_N * dev::SM_SEG : 0));
                DevBuffer big;
                int32_t* bd = static_cast<int32_t*>(big.get(size_t(nruns) * rbytes, s));
                dev::submesh_batch_run_emit(d_items, d_pp, PW, n, sets, offsets, bd, nruns, s);
                hr   = ctx.readback->acquire(gb + size_t(nruns) * rbytes);
                hs   = reinterpret_cast<int32_t*>(hr);
                runs = reinterpret_cast<int32_t*>(hr + gb);
                hip_check(hipMemcpyAsync(hs, dseg, gb, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(submesh runs)");
                hip_check(hipMemcpyAsync(runs, bd, size_t(nruns) * rbytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(submesh runs)");
                ctx.sync();
            }
            td.reset();
            {
                ScopedTimer tc(nullptr, "mesh completion");
                std::vector<const MeshConfig*> cfgs(static_cast<size_t>(n));
                std::vector<MRMeshData*> out(static_cast<size_t>(n));
                for (int q = 0; q < n; ++q)
                {
                    const size_t j  = size_t(sel[size_t(q)]);
                    cfgs[size_t(q)] = &ms[j]->host.cfg;
                    out[size_t(q)]  = &meshes[j];
                }
                meshes_from_runs(cfgs, runs, hs, out);
            }
            if (with_lists)
            {
                // the lists' sizes and levels, then their runs
                ScopedTimer te(&ctx, "mesh ids on device");
                if (table.empty())
                {
                    ctx.sync();
                    table.assign(ht, ht + size_t(n) * dev::DL_N * dev::SM_SEG);
                }
                emit_batch_lists(ctx, d_items, d_cp, lw, n, lbits, lscan, table, sel, *lists, *list_mem);
            }
        }

        // update_cell_array_from_tag and the make_graduation fixed point of
        // the meshes ms (tags[j] indexed by mesh j's storage) on the device,
        // all meshes in each launch (rebuild_batch.hip).  Returns, per mesh,
        // whether its leaves change, and the new leaves of those that do; with
        // meshes, the changed meshes are built on the device too
        // (construct_batch) and meshes[j] is the complete new mesh of mesh j
        // (out is then left empty); with lists, their work lists too.
        std::vector<bool> device_update_batch(Ctx& ctx, const std::vector<const Mesh*>& ms, const std::vector<const uint8_t*>& tags,
                                              std::vector<CellArray>& out, std::vector<MRMeshData>* meshes,
                                              std::vector<std::array<DeviceLists::One, dev::DL_N>>* lists = nullptr,
                                              std::shared_ptr<DevBuffer>* list_mem = nullptr)
        {
            hipStream_t s = ctx.stream;
            const int n   = int(ms.size());
            std::vector<dev::MeshItem> items(static_cast<size_t>(n));
            std::vector<int64_t> wpref(size_t(n) + 1, 0), cpref(size_t(n) + 1, 0), rpref(size_t(n) + 1, 0);
            int w2 = -1;
            for (int j = 0; j < n; ++j)
            {
                const Mesh& m       = *ms[size_t(j)];
                const MeshConfig& c = m.host.cfg;
                const dev::Bits b   = make_bits(c);
                if (w2 >= 0 && w2 != 2 * c.grad_width)
                {
                    throw std::invalid_argument("batched mesh update: instances of different graduation widths");
                }
                w2                   = 2 * c.grad_width;
                dev::MeshItem& it    = items[size_t(j)];
                it.b                 = b;
                it.cells             = m.dev->cells.d;
                it.tag               = tags[size_t(j)];
                it.w0                = wpref[size_t(j)];
                it.dim               = c.dim;
                it.min_level         = c.min_level;
                it.max_level         = c.max_level;
                wpref[size_t(j) + 1] = wpref[size_t(j)] + b.off[b.hi + 1];
                cpref[size_t(j) + 1] = cpref[size_t(j)] + it.cells.ncells;
                rpref[size_t(j) + 1] = rpref[size_t(j)] + it.cells.nruns;
            }
            const int64_t W       = wpref[size_t(n)];
            constexpr int kPasses = 3;
            const int fs          = kPasses + 1;
            // the item table and its prefixes, one upload
            const size_t ib = (size_t(n) * sizeof(dev::MeshItem) + 255) & ~size_t(255);
            const size_t pb = size_t(n + 1) * 8;
            const size_t hb = ib + 3 * pb;
            char* h         = ctx.staging->acquire(hb);
            std::memcpy(h, items.data(), size_t(n) * sizeof(dev::MeshItem));
            std::memcpy(h + ib, wpref.data(), pb);
            std::memcpy(h + ib + pb, cpref.data(), pb);
            std::memcpy(h + ib + 2 * pb, rpref.data(), pb);
            DevBuffer tab;
            char* dt = static_cast<char*>(tab.get(hb, s));
            hip_check(hipMemcpyAsync(dt, h, hb, hipMemcpyHostToDevice, s), "hipMemcpyAsync(mesh items)");
            ctx.staging->release(s);
            const auto* d_items = reinterpret_cast<const dev::MeshItem*>(dt);
            const auto* d_wpref = reinterpret_cast<const int64_t*>(dt + ib);
            const auto* d_cpref = reinterpret_cast<const int64_t*>(dt + ib + pb);
            const auto* d_rpref = reinterpret_cast<const int64_t*>(dt + ib + 2 * pb);
            // bitmaps: old, add, rem, two pass buffers, the projected marks; flags
            DevBuffer work;
            uint64_t* w = static_cast<uint64_t*>(work.get(size_t(6 * std::max<int64_t>(W, 1)) * 8 + size_t(n) * fs * 4 + 256, s));
            uint64_t* old = w, *add = w + W, *rem = w + 2 * W, *q = w + 5 * W;
            uint64_t* buf[2] = {w + 3 * W, w + 4 * W};
            int* flags       = reinterpret_cast<int*>(w + 6 * W);
            dev::mesh_batch_zero(w, 3 * W, s);
            dev::mesh_batch_leaves(d_items, d_rpref, rpref[size_t(n)], n, old, s);
            dev::mesh_batch_tag_marks(d_items, d_cpref, cpref[size_t(n)], n, add, rem, s);
            dev::mesh_batch_combine(old, add, rem, buf[0], W, s);
            int cur = 0;
            std::vector<int> hf(size_t(n) * fs);
            for (int batch = 0;; ++batch)
            {
                if (batch > 21)
                {
                    throw std::logic_error("make_graduation: no fixed point");
                }
                hip_check(hipMemsetAsync(flags, 0, size_t(n) * fs * 4, s), "hipMemsetAsync(flags)");
                for (int p = 0; p < kPasses; ++p)
                {
                    dev::mesh_batch_graduation_pass(d_items, d_wpref, W, n, buf[cur], q, buf[1 - cur], w2, flags, fs, p, s);
                    cur = 1 - cur;
                }
                dev::mesh_batch_differ(d_wpref, W, n, buf[cur], old, flags, fs, kPasses, s);
                hip_check(hipMemcpyAsync(hf.data(), flags, hf.size() * 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(flags)");
                ctx.sync();
                bool all = true;
                for (int j = 0; j < n && all; ++j)
                {
                    bool fixed = false;
                    for (int p = 0; p < kPasses; ++p)
                    {
                        fixed = fixed || hf[size_t(j) * fs + p] == 0;
                    }
                    all = fixed;
                }
                if (all)
                {
                    break;
                }
            }
            std::vector<bool> changed(static_cast<size_t>(n));
            int64_t cap = 1024;
            for (int j = 0; j < n; ++j)
            {
                changed[size_t(j)] = hf[size_t(j) * fs + kPasses] != 0;
                if (changed[size_t(j)])
                {
                    cap += items[size_t(j)].cells.nruns + items[size_t(j)].cells.nruns / 4;
                }
            }
            out.assign(size_t(n), CellArray());
            if (cap == 1024 && std::none_of(changed.begin(), changed.end(), [](bool c) { return c; }))
            {
                return changed;
            }
            if (meshes)
            {
                meshes->assign(size_t(n), MRMeshData());
                if (lists)
                {
                    lists->assign(size_t(n), {});
                }
                construct_batch(ctx, ms, items, changed, buf[cur], *meshes, old, lists, list_mem);
                return changed;
            }
            // the new leaves of the changed meshes as runs, one round trip
            const size_t tbytes = dev::mesh_batch_scan_temp_bytes(W);
            const size_t cbytes = (size_t(2 * std::max<int64_t>(W, 1)) * 4 + 255) & ~size_t(255);
            DevBuffer scan;
            char* sp         = static_cast<char*>(scan.get(cbytes + tbytes + size_t(cap) * 24, s));
            int32_t* counts  = reinterpret_cast<int32_t*>(sp);
            int32_t* offsets = counts + std::max<int64_t>(W, 1);
            void* temp       = sp + cbytes;
            int32_t* dr      = reinterpret_cast<int32_t*>(sp + cbytes + tbytes);
            const uint64_t* leaves = buf[cur];
            dev::mesh_batch_run_offsets(d_items, d_wpref, W, n, leaves, flags, fs, kPasses, counts, offsets, temp, tbytes, s);
            dev::mesh_batch_run_emit(d_items, d_wpref, W, n, leaves, flags, fs, kPasses, offsets, dr, cap, s);
            char* hr         = ctx.readback->acquire(8 + size_t(cap) * 24);
            int32_t* last    = reinterpret_cast<int32_t*>(hr);
            int32_t* runs    = reinterpret_cast<int32_t*>(hr + 8);
            hip_check(hipMemcpyAsync(&last[0], offsets + W - 1, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(runs)");
            hip_check(hipMemcpyAsync(&last[1], counts + W - 1, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(runs)");
            hip_check(hipMemcpyAsync(runs, dr, size_t(cap) * 24, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(runs)");
            ctx.sync();
            const int64_t nruns = int64_t(last[0]) + last[1];
            if (nruns > cap)
            {
                DevBuffer big;
                int32_t* bd = static_cast<int32_t*>(big.get(size_t(nruns) * 24, s));
                dev::mesh_batch_run_emit(d_items, d_wpref, W, n, leaves, flags, fs, kPasses, offsets, bd, nruns, s);
                hr   = ctx.readback->acquire(8 + size_t(nruns) * 24);
                runs = reinterpret_cast<int32_t*>(hr + 8);
                hip_check(hipMemcpyAsync(runs, bd, size_t(nruns) * 24, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(runs)");
                ctx.sync();
            }
            // the runs come in mesh order: each changed mesh is one segment
            auto first = [&](int j)
            {
                int64_t lo = 0, hi = nruns;
                while (lo < hi)
                {
                    const int64_t mid = (lo + hi) >> 1;
                    if (runs[6 * mid + 5] < j)
                    {
                        lo = mid + 1;
                    }
                    else
                    {
                        hi = mid;
                    }
                }
                return lo;
            };
            parallel_for(n,
                         [&](int j)
                         {
                             if (!changed[size_t(j)])
                             {
                                 return;
                             }
                             CellArray& ca = out[size_t(j)];
                             ca            = empty_cell_array(items[size_t(j)].dim);
                             for (int64_t k = first(j), e = first(j + 1); k < e; ++k)
                             {
                                 const int32_t* r = runs + 6 * k;
                                 ca[size_t(r[4])].push(r[2], r[3], r[0], r[1]);
                             }
                         });
            return changed;
        }
    } // namespace

    // ------------------------------------------------------------------
    // ProgramBuf
    // ------------------------------------------------------------------
    ProgramBuf::~ProgramBuf()
    {
        drop_graph();
        if (copied_)
        {
            (void)hipEventSynchronize(copied_);
            (void)hipEventDestroy(copied_);
        }
        if (hbuf_)
        {
            (void)hipHostFree(hbuf_);
        }
        if (dbuf_)
        {
            dev_free(dbuf_, ctx_->stream);
        }
    }

    void ProgramBuf::drop_graph()
    {
        if (gexec_)
        {
            // a launched replay may still be running
            (void)hipStreamSynchronize(ctx_->stream);
            (void)hipGraphExecDestroy(gexec_);
            gexec_ = nullptr;
        }
    }

    void ProgramBuf::clear()
    {
        segs_.clear();
        seg_pref_.assign(1, 0);
        stage_seg_.assign(1, 0);
        stage_cells_.clear();
        stage_max_chain_.clear();
        chain_items_.clear();
        nchains_           = -1;
        chains_consistent_ = true;
        plan_.clear();
        blk_.clear();
        witems_.clear();
        stage_item_.assign(1, 0);
        stage_nitems_.clear();
        aux_.clear();
        uploaded_ = false;
    }

    void ProgramBuf::append(const std::vector<Chain>& chains)
    {
        if (nchains_ < 0)
        {
            nchains_ = int(chains.size());
        }
        else if (nchains_ != int(chains.size()))
        {
            chains_consistent_ = false;
        }
        if (chain_items_.size() < chains.size())
        {
            chain_items_.resize(chains.size());
        }
        size_t depth = 0;
        for (const Chain& c : chains)
        {
            depth = std::max(depth, c.size());
        }
        for (size_t k = 0; k < depth; ++k)
        {
            int64_t cells = 0, most = 0;
            const int32_t stage = int32_t(stage_cells_.size());
            for (size_t i = 0; i < chains.size(); ++i)
            {
                const Chain& c = chains[i];
                if (k >= c.size())
                {
                    continue;
                }
                int64_t mine = 0;
                for (const dev::Segment& g : c[k])
                {
                    if (g.c1 <= g.c0)
                    {
                        continue;
                    }
                    check_segment(g);
                    chain_items_[i].emplace_back(int32_t(segs_.size()), stage);
                    segs_.push_back(g);
                    seg_pref_.push_back(seg_pref_.back() + (g.c1 - g.c0));
                    cells += g.c1 - g.c0;
                    mine += g.c1 - g.c0;
                }
                most = std::max(most, mine);
            }
            if (cells > 0)
            {
                stage_seg_.push_back(int32_t(segs_.size()));
                stage_cells_.push_back(cells);
                stage_max_chain_.push_back(most);
            }
        }
        uploaded_ = false;
    }

    // consecutive stages where no chain has more than block_stage_max() cells
    // become one block launch; every other stage is a grid-wide launch
    void ProgramBuf::build_block_plan()
    {
        plan_.clear();
        blk_.clear();
        if (!blocks_capable())
        {
            return;
        }
        const int ns        = nstages();
        const int64_t limit = block_stage_max();
        const size_t nch    = size_t(nchains_);
        std::vector<std::vector<int32_t>> per(nch); // per block step and chain
        std::vector<size_t> cur(nch, 0);
        std::vector<int32_t> items;
        std::vector<int32_t> offs;
        for (int k = 0; k < ns;)
        {
            if (stage_max_chain_[size_t(k)] > limit)
            {
                plan_.push_back({false, k, k + 1, 0});
                for (size_t i = 0; i < nch; ++i)
                {
                    while (cur[i] < chain_items_[i].size() && chain_items_[i][cur[i]].second == k)
                    {
                        ++cur[i];
                    }
                }
                ++k;
                continue;
            }
            int k1 = k;
            while (k1 < ns && stage_max_chain_[size_t(k1)] <= limit)
            {
                ++k1;
            }
            plan_.push_back({true, k, k1, offs.size()});
            for (size_t i = 0; i < nch; ++i)
            {
                offs.push_back(int32_t(items.size()));
                const auto& ci = chain_items_[i];
                while (cur[i] < ci.size() && ci[cur[i]].second < k1)
                {
                    const auto [seg, stage] = ci[cur[i]];
                    ++cur[i];
                    // a barrier when the chain's next item in this launch is of a later stage
                    const bool barrier = cur[i] < ci.size() && ci[cur[i]].second < k1 && ci[cur[i]].second != stage;
                    items.push_back(int32_t(seg) << 1 | (barrier ? 1 : 0));
                }
            }
            offs.push_back(int32_t(items.size()));
            k = k1;
        }
        // chain offsets are relative to the items: shift them past the offsets
        blk_items_ = offs.size();
        blk_       = std::move(offs);
        for (int32_t& o : blk_)
        {
            o += int32_t(blk_items_);
        }
        blk_.insert(blk_.end(), items.begin(), items.end());
    }

    void ProgramBuf::launch_block_plan(const Tuning& t, hipStream_t s) const
    {
        for (const PlanStep& st : plan_)
        {
            if (st.blocks)
            {
                dev::run_program_blocks(prog_, dblk_ + st.off, dblk_, nchains_, kBlockThreads, s);
            }
            else
            {
                launch_stage(t, st.k0, s);
            }
        }
    }

    void ProgramBuf::launch_stage(const Tuning& t, int k, hipStream_t s) const
    {
        if (t.cpt <= 0)
        {
            dev::run_program_stage(prog_, k, stage_nitems_[size_t(k)], s);
        }
        else
        {
            dev::run_program_stage_cells(prog_, k, stage_cells_[size_t(k)], t.block, t.cpt, s);
        }
    }

    void ProgramBuf::launch_stagewise(const Tuning& t, hipStream_t s) const
    {
        for (int k = 0; k < nstages(); ++k)
        {
            launch_stage(t, k, s);
        }
    }

    // the segments of a stage packed densely in wave items of 64 consecutive
    // indices, each with the segment its first index falls in; a large
    // segment of a streaming operator gets items of kStreamChunks chunks
    void ProgramBuf::build_wave_items()
    {
        constexpr int kStreamChunks = 16;
        witems_.clear();
        stage_item_.assign(1, 0);
        stage_nitems_.clear();
        auto streaming = [&](int32_t sg)
        {
            const dev::Segment& g = segs_[size_t(sg)];
            return dev::is_streaming_op(g.op) && g.c1 - g.c0 >= 64 * kStreamChunks;
        };
        for (size_t k = 0; k + 1 < stage_seg_.size(); ++k)
        {
            const int32_t se = stage_seg_[k + 1];
            for (int32_t sb = stage_seg_[k]; sb < se;)
            {
                if (streaming(sb))
                {
                    const int64_t n = segs_[size_t(sb)].c1 - segs_[size_t(sb)].c0;
                    for (int64_t off = 0; off < n; off += 64 * kStreamChunks)
                    {
                        if (off > int64_t(INT32_MAX))
                        {
                            throw std::length_error("stage program: segment too large");
                        }
                        witems_.push_back({sb, int32_t(off), sb + 1, kStreamChunks});
                    }
                    ++sb;
                    continue;
                }
                // a pack: the consecutive segments [sb, end) that are not streamed
                int32_t end = sb + 1;
                while (end < se && !streaming(end))
                {
                    ++end;
                }
                int32_t cur = sb;
                for (int64_t pos = seg_pref_[size_t(sb)]; pos < seg_pref_[size_t(end)]; pos += 64)
                {
                    while (seg_pref_[size_t(cur) + 1] <= pos)
                    {
                        ++cur;
                    }
                    const int64_t off = pos - seg_pref_[size_t(cur)];
                    if (off > int64_t(INT32_MAX))
                    {
                        throw std::length_error("stage program: segment too large");
                    }
                    witems_.push_back({cur, int32_t(off), end, 1});
                }
                sb = end;
            }
            stage_nitems_.push_back(int64_t(witems_.size()) - stage_item_.back());
            stage_item_.push_back(int64_t(witems_.size()));
        }
    }

    size_t ProgramBuf::add_aux(const void* src, size_t n)
    {
        const size_t off = align_up(aux_.size(), 16);
        aux_.resize(off + n);
        std::memcpy(aux_.data() + off, src, n);
        return off;
    }

    int64_t ProgramBuf::max_stage_cells() const
    {
        int64_t m = 0;
        for (int64_t c : stage_cells_)
        {
            m = std::max(m, c);
        }
        return m;
    }

    int64_t ProgramBuf::total_cells() const
    {
        int64_t t = 0;
        for (int64_t c : stage_cells_)
        {
            t += c;
        }
        return t;
    }

    void ProgramBuf::upload()
    {
        upload_with(true, true);
    }

    void ProgramBuf::upload_for(const Tuning& t)
    {
        upload_with(t.cpt == 0 && t.mode != MODE_FUSED, t.mode == MODE_BLOCKS);
    }

    void ProgramBuf::upload_with(bool items, bool plan)
    {
        hipStream_t s        = ctx_->stream;
        const size_t nseg    = segs_.size();
        const size_t o_segs  = 0;
        const size_t o_pref  = align_up(o_segs + nseg * sizeof(dev::Segment), 256);
        const size_t o_stage = align_up(o_pref + (nseg + 1) * sizeof(int64_t), 256);
        const size_t o_aux   = align_up(o_stage + stage_seg_.size() * sizeof(int32_t), 256);
        if (plan)
        {
            build_block_plan();
        }
        else
        {
            plan_.clear();
            blk_.clear();
        }
        if (items)
        {
            build_wave_items();
        }
        else
        {
            witems_.clear();
            stage_item_.assign(1, 0);
            stage_nitems_.clear();
        }
        items_built_ = items;
        plan_built_  = plan;
        const size_t o_blk   = align_up(o_aux + aux_.size(), 256);
        const size_t o_sit   = align_up(o_blk + blk_.size() * sizeof(int32_t), 256);
        const size_t o_wit   = align_up(o_sit + stage_item_.size() * sizeof(int64_t), 256);
        const size_t total   = std::max<size_t>(o_wit + witems_.size() * sizeof(dev::WaveItem), 256);
        drop_graph();
        if (!copied_)
        {
            hip_check(hipEventCreateWithFlags(&copied_, hipEventDisableTiming), "hipEventCreate");
        }
        else
        {
            // the staging buffer is reused: the previous copy must have read it
            hip_check(hipEventSynchronize(copied_), "hipEventSynchronize");
        }
        if (total > dcap_)
        {
            if (dbuf_)
            {
                dev_free(dbuf_, s);
            }
            dcap_ = std::max(total, 2 * dcap_);
            dbuf_ = static_cast<char*>(dev_alloc(dcap_, s));
        }
        if (total > hcap_)
        {
            if (hbuf_)
            {
                hip_check(hipHostFree(hbuf_), "hipHostFree(program)");
                hbuf_ = nullptr;
            }
            hcap_ = std::max(total, 2 * hcap_);
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&hbuf_), hcap_, hipHostMallocDefault), "hipHostMalloc(program)");
        }
        auto* hsegs = reinterpret_cast<dev::Segment*>(hbuf_ + o_segs);
        std::memcpy(hsegs, segs_.data(), nseg * sizeof(dev::Segment));
        for (size_t i = 0; i < nseg; ++i)
        {
            dev::Segment& g = hsegs[i];
            if (g.aux & AUX_THR)
            {
                g.thr = reinterpret_cast<const double*>(dbuf_ + o_aux + reinterpret_cast<uintptr_t>(g.thr));
            }
            if (g.aux & AUX_TABS)
            {
                g.tabs = reinterpret_cast<const dev::Table*>(dbuf_ + o_aux + reinterpret_cast<uintptr_t>(g.tabs));
            }
            g.aux = 0;
        }
        std::memcpy(hbuf_ + o_pref, seg_pref_.data(), (nseg + 1) * sizeof(int64_t));
        std::memcpy(hbuf_ + o_stage, stage_seg_.data(), stage_seg_.size() * sizeof(int32_t));
        if (!aux_.empty())
        {
            std::memcpy(hbuf_ + o_aux, aux_.data(), aux_.size());
        }
        if (!blk_.empty())
        {
            std::memcpy(hbuf_ + o_blk, blk_.data(), blk_.size() * sizeof(int32_t));
        }
        std::memcpy(hbuf_ + o_sit, stage_item_.data(), stage_item_.size() * sizeof(int64_t));
        if (!witems_.empty())
        {
            std::memcpy(hbuf_ + o_wit, witems_.data(), witems_.size() * sizeof(dev::WaveItem));
        }
        hip_check(hipMemcpyAsync(dbuf_, hbuf_, total, hipMemcpyHostToDevice, s), "hipMemcpyAsync(program)");
        hip_check(hipEventRecord(copied_, s), "hipEventRecord");
        prog_.segs      = reinterpret_cast<const dev::Segment*>(dbuf_ + o_segs);
        prog_.seg_pref  = reinterpret_cast<const int64_t*>(dbuf_ + o_pref);
        prog_.stage_seg = reinterpret_cast<const int32_t*>(dbuf_ + o_stage);
        prog_.nstages   = int(stage_cells_.size());
        dblk_           = reinterpret_cast<const int32_t*>(dbuf_ + o_blk);
        prog_.stage_item = reinterpret_cast<const int64_t*>(dbuf_ + o_sit);
        prog_.items      = reinterpret_cast<const dev::WaveItem*>(dbuf_ + o_wit);
        uploaded_       = true;
    }

    void ProgramBuf::run(const Tuning& t)
    {
        if (!uploaded_)
        {
            throw std::logic_error("stage program run before upload");
        }
        if (prog_.nstages == 0)
        {
            return;
        }
        hipStream_t s = ctx_->stream;
        if (t.mode == MODE_FUSED)
        {
            // the grid barrier costs grow with the number of blocks that arrive
            // at it: at most one block per CU, the rest is grid-strided
            // the fused launch walks cells: cpt 0 (wave items) walks one per thread
            const int cpt         = std::max(t.cpt, 1);
            const int maxb        = std::min(max_blocks_cached(t.block), cu_count());
            const int64_t per     = int64_t(t.block) * cpt;
            const int64_t needed  = (max_stage_cells() + per - 1) / per;
            const int nblocks     = int(std::max<int64_t>(1, std::min<int64_t>(needed, maxb)));
            if (maxb > 0 && dev::run_program_fused(prog_, t.block, nblocks, cpt, s))
            {
                return;
            }
            // the runtime refused the cooperative launch: same program, one launch per stage
            (void)hipGetLastError();
            launch_stagewise(t, s);
            return;
        }
        if (t.cpt == 0 && t.mode != MODE_FUSED && !items_built_)
        {
            throw std::logic_error("stage program uploaded without the wave items its launch shape uses");
        }
        if (t.mode == MODE_GRAPH || t.mode == MODE_BLOCKS)
        {
            const bool blocks = t.mode == MODE_BLOCKS && plan_built_ && blocks_capable();
            const int gmode   = blocks ? MODE_BLOCKS : MODE_GRAPH;
            if (!gexec_ || g_block_ != t.block || g_cpt_ != t.cpt || g_mode_ != gmode)
            {
                drop_graph();
                hipGraph_t g = nullptr;
                hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                if (blocks)
                {
                    launch_block_plan(t, s);
                }
                else
                {
                    launch_stagewise(t, s);
                }
                hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
                const hipError_t e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                hip_check(e, "hipGraphInstantiate");
                g_block_ = t.block;
                g_cpt_   = t.cpt;
                g_mode_  = gmode;
            }
            hip_check(hipGraphLaunch(gexec_, s), "hipGraphLaunch");
            return;
        }
        launch_stagewise(t, s);
    }

    // ------------------------------------------------------------------
    // Batch
    // ------------------------------------------------------------------
    Batch::Batch(Ctx* ctx, std::vector<BatchInstance> inst)
        : ctx_(ctx), inst_(std::move(inst)), adapt_prog_(ctx), resolve_prog_(ctx), uf_prog_(ctx)
    {
        ctx_->require_device();
        std::vector<const Mesh*> meshes;
        for (const BatchInstance& b : inst_)
        {
            if (!b.u || !b.unp1 || b.u == b.unp1)
            {
                throw std::invalid_argument("batch instance needs two distinct fields u and unp1");
            }
            if (b.u->mesh != b.unp1->mesh)
            {
                throw std::invalid_argument("batch instance fields u and unp1 are on different meshes");
            }
            if (b.u->mesh->ctx != ctx_)
            {
                throw std::invalid_argument("batch instance mesh belongs to another context");
            }
            meshes.push_back(b.u->mesh);
        }
        std::sort(meshes.begin(), meshes.end());
        if (std::adjacent_find(meshes.begin(), meshes.end()) != meshes.end())
        {
            throw std::invalid_argument("two batch instances share a mesh");
        }
        fv_prog_[0] = std::make_unique<ProgramBuf>(ctx_);
        fv_prog_[1] = std::make_unique<ProgramBuf>(ctx_);
    }

    Batch::~Batch()
    {
        release_replaced();
        hipStream_t s = ctx_->stream;
        if (det_arena_)
        {
            dev_free(det_arena_, s);
        }
        if (tag_arena_)
        {
            dev_free(tag_arena_, s);
        }
        if (flags_)
        {
            dev_free(flags_, s);
        }
        (void)hipStreamSynchronize(s);
        if (host_tags_)
        {
            (void)hipHostFree(host_tags_);
        }
        if (host_flags_)
        {
            (void)hipHostFree(host_flags_);
        }
    }

    void Batch::check_instances()
    {
        for (BatchInstance& b : inst_)
        {
            b.u->check_current();
            // unp1 is the step's output: resized (zero-filled) here when stale, as
            // the demo loop does with unp1.resize() (advection_2d.cpp:143)
            if (b.unp1->version != b.unp1->mesh->version || !b.unp1->d || b.unp1->size != b.unp1->mesh->host.nref)
            {
                b.unp1->allocate(true);
            }
        }
    }

    int64_t Batch::nb_cells() const
    {
        int64_t n = 0;
        for (const BatchInstance& b : inst_)
        {
            n += b.u->mesh->host.nb_cells(CELLS);
        }
        return n;
    }

    void Batch::set_tuning(const Tuning& t)
    {
        if (t.block < 64 || t.block > 1024 || (t.block & 63) != 0)
        {
            throw std::invalid_argument("block size must be a multiple of 64 in [64, 1024]");
        }
        if (t.cpt < 0 || t.cpt > 64)
        {
            throw std::invalid_argument("cells per thread must be in [0, 64] (0: wave items)");
        }
        if (t.mode < MODE_FUSED || t.mode > MODE_BLOCKS)
        {
            throw std::invalid_argument("unknown program mode");
        }
        tuning_ = t;
    }

    void Batch::ensure_images(const std::vector<int>& ids)
    {
        std::vector<int> need;
        for (int i : ids)
        {
            const Mesh& m = *inst_[size_t(i)].u->mesh;
            if (!m.dev || !m.dev->has_ghost || !m.dev->has_adapt)
            {
                need.push_back(i);
            }
        }
        if (need.empty())
        {
            return;
        }
        std::vector<HostImage> imgs(need.size());
        parallel_for(int(need.size()), [&](int k) { imgs[size_t(k)] = build_host_image(inst_[size_t(need[size_t(k)])].u->mesh->host); });
        for (size_t k = 0; k < need.size(); ++k)
        {
            Mesh& m = *inst_[size_t(need[k])].u->mesh;
            m.dev   = upload_image(*ctx_, m.host, std::move(imgs[k]), false);
        }
        ctx_->check_resolve();
    }

    void Batch::ensure_arenas(int64_t n)
    {
        hipStream_t s = ctx_->stream;
        if (!flags_)
        {
            flags_ = static_cast<int*>(dev_alloc(std::max<size_t>(inst_.size(), 1) * sizeof(int), s));
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_flags_), std::max<size_t>(inst_.size(), 1) * sizeof(int),
                                    hipHostMallocDefault),
                      "hipHostMalloc(flags)");
        }
        if (n > arena_cap_)
        {
            if (det_arena_)
            {
                dev_free(det_arena_, s);
                dev_free(tag_arena_, s);
            }
            arena_cap_ = std::max<int64_t>(n, arena_cap_ + arena_cap_ / 2);
            det_arena_ = static_cast<double*>(dev_alloc(size_t(arena_cap_) * 8, s));
            tag_arena_ = static_cast<uint8_t*>(dev_alloc(size_t(arena_cap_), s));
        }
        if (n > host_cap_)
        {
            if (host_tags_)
            {
                hip_check(hipHostFree(host_tags_), "hipHostFree(tags)");
                host_tags_ = nullptr;
            }
            host_cap_ = std::max<int64_t>(n, host_cap_ + host_cap_ / 2);
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_tags_), size_t(host_cap_), hipHostMallocDefault), "hipHostMalloc(tags)");
        }
    }

    int Batch::adapt(double eps, double regularity)
    {
        check_instances();
        ScopedTimer timer(ctx_, "mesh adaptation");
        hipStream_t s = ctx_->stream;
        std::vector<int> act;
        for (size_t i = 0; i < inst_.size(); ++i)
        {
            const MeshConfig& c = inst_[i].u->mesh->host.cfg;
            if (c.min_level < c.max_level)
            {
                act.push_back(int(i));
            }
        }
        int calls = 0;
        for (int ite = 0;; ++ite)
        {
            act.erase(std::remove_if(act.begin(), act.end(),
                                     [&](int i)
                                     {
                                         const MeshConfig& c = inst_[size_t(i)].u->mesh->host.cfg;
                                         return ite >= c.max_level - c.min_level;
                                     }),
                      act.end());
            if (act.empty())
            {
                break;
            }
            ensure_images(act);
            const size_t na = act.size();
            std::vector<int64_t> off(na + 1, 0);
            for (size_t k = 0; k < na; ++k)
            {
                off[k + 1] = off[k] + int64_t(align_up(size_t(inst_[size_t(act[k])].u->mesh->host.nref), 256));
            }
            ensure_arenas(std::max<int64_t>(off[na], 1));

            // harten (mr/adapt.hpp:276-389) on the device: tags, ghosts, details, criteria, maximum
            {
                ScopedTimer t(ctx_, "harten device pass");
                std::optional<ScopedTimer> tb(std::in_place, nullptr, "harten program build");
                adapt_prog_.clear();
                std::vector<Chain> chains(na);
                // the criteria thresholds of every instance (the program's aux
                // data), then the instances' chains in parallel
                std::vector<size_t> thr_at(na);
                for (size_t k = 0; k < na; ++k)
                {
                    const dev::Thresholds thr = make_thresholds(inst_[size_t(act[k])].u->mesh->host.cfg, ite, eps, regularity);
                    thr_at[k]                 = adapt_prog_.add_aux(&thr, sizeof(thr));
                }
                parallel_for(int(na), [&](int kk)
                {
                    const size_t k      = size_t(kk);
                    Field& f            = *inst_[size_t(act[k])].u;
                    Mesh& m             = *f.mesh;
                    DevMesh& dm         = *m.dev;
                    const MeshConfig& c = m.host.cfg;
                    const int64_t n     = m.host.nref;
                    double* det         = det_arena_ + off[k];
                    uint8_t* tag        = tag_arena_ + off[k];
                    Chain pre(2);
                    dev::Segment fd = make_seg(dev::SEG_FILL_F64, c);
                    fd.out          = det;
                    fd.c1           = n;
                    dev::Segment ft = make_seg(dev::SEG_FILL_U8, c);
                    ft.tag          = tag;
                    ft.c1           = n;
                    dev::Segment ff = make_seg(dev::SEG_FILL_U8, c);
                    ff.tag          = reinterpret_cast<uint8_t*>(flags_ + k);
                    ff.c1           = sizeof(int);
                    pre[0]          = {fd, ft, ff};
                    dev::Segment ti = list_seg(dev::SEG_TAG_INIT, c, dm.cells, 0, dm.cells.d.ncells);
                    ti.tag          = tag;
                    pre[1]          = {ti};
                    Chain g;
                    ghost_chain(g, dm, c, f.d, f.has_bc, f.bc_value);
                    Chain ch = zip(pre, g);
                    int64_t c0, c1;
                    dm.detail.cells_of_levels(c.min_level > 0 ? c.min_level - 1 : 0, c.max_level - ite, c0, c1);
                    dev::Segment sd = list_seg(dev::SEG_DETAIL, c, dm.detail, c0, c1);
                    sd.u            = f.d;
                    sd.out          = det;
                    push_stage(ch, sd);
                    const size_t h = thr_at[k];
                    dm.criteria.cells_of_levels(c.min_level - 1, c.max_level - ite, c0, c1);
                    dev::Segment sc = list_seg(dev::SEG_CRITERIA, c, dm.criteria, c0, c1);
                    sc.out          = det;
                    sc.tag          = tag;
                    sc.thr          = reinterpret_cast<const double*>(uintptr_t(h));
                    sc.aux          = AUX_THR;
                    push_stage(ch, sc);
                    for (int level = c.max_level; level > 0; --level)
                    {
                        dm.maximum.cells_of_levels(level - 1, level, c0, c1);
                        dev::Segment sm = list_seg(dev::SEG_MAXIMUM, c, dm.maximum, c0, c1);
                        sm.tag          = tag;
                        push_stage(ch, sm);
                    }
                    dev::Segment tc = list_seg(dev::SEG_TAG_CHANGE, c, dm.cells, 0, dm.cells.d.ncells);
                    tc.tag          = tag;
                    tc.err          = flags_ + k;
                    push_stage(ch, tc);
                    chains[k] = std::move(ch);
                });
                adapt_prog_.append(chains);
                adapt_prog_.upload_for(pass_tuning());
                tb.reset();
                adapt_prog_.run(pass_tuning());
                calls += int(na);
                // the versions the previous pass replaced, freed on the host
                // while this pass runs
                release_replaced();
            }

            // tags to the host, then every instance's mesh update in parallel
            std::vector<Rebuild> rb(na);
            // the arrays of the device-built lists of this pass's new meshes
            std::shared_ptr<DevBuffer> batch_list_mem;
            {
                ScopedTimer t(ctx_, "mesh update");
                std::optional<ScopedTimer> tt(std::in_place, ctx_, "tags to host");
                // which instances' tags change their cell array (graduated meshes
                // whose tags change nothing are at the fixed point already)
                hip_check(hipMemcpyAsync(host_flags_, flags_, na * sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(flags)");
                ctx_->sync();
                std::vector<size_t> work;
                for (size_t k = 0; k < na; ++k)
                {
                    const Mesh& m = *inst_[size_t(act[k])].u->mesh;
                    if (host_flags_[k] == 0 && m.host.graduated)
                    {
                        rb[k].converged = true;
                        continue;
                    }
                    work.push_back(k);
                }
                if (!batch_host_rebuild())
                {
                    // the mesh update and the changed meshes' ids on the device,
                    // all instances per launch; the host completes those meshes
                    // (corners, renumbering) and builds their programs
                    // (SMR_HOST_SUBMESH=1: their ids on the host)
                    tt.reset();
                    std::vector<const Mesh*> ms;
                    std::vector<const uint8_t*> tg;
                    for (size_t k : work)
                    {
                        ms.push_back(inst_[size_t(act[k])].u->mesh);
                        tg.push_back(tag_arena_ + off[k]);
                    }
                    std::vector<CellArray> cas;
                    std::vector<MRMeshData> nms;
                    std::vector<std::array<DeviceLists::One, dev::DL_N>> dls;
                    std::vector<bool> ch;
                    const bool dev_ids   = !host_submesh();
                    const bool dev_lists = dev_ids && !host_lists();
                    {
                        ScopedTimer td(ctx_, "device mesh update");
                        ch = ms.empty() ? std::vector<bool>()
                                        : device_update_batch(*ctx_, ms, tg, cas, dev_ids ? &nms : nullptr, dev_lists ? &dls : nullptr,
                                                              &batch_list_mem);
                    }
                    ScopedTimer th(ctx_, "host mesh and programs");
                    parallel_for(int(work.size()),
                                 [&](int q)
                                 {
                                     const size_t k      = work[size_t(q)];
                                     const Mesh& m       = *inst_[size_t(act[k])].u->mesh;
                                     const MeshConfig& c = m.host.cfg;
                                     Rebuild& r          = rb[k];
                                     if (!ch[size_t(q)])
                                     {
                                         r.converged = true;
                                         return;
                                     }
                                     r.nm           = dev_ids ? std::move(nms[size_t(q)]) : build_mesh(c, std::move(cas[size_t(q)]));
                                     r.nm.graduated = true;
                                     if (dev_lists)
                                     {
                                         r.device_lists = true;
                                         r.dl           = std::move(dls[size_t(q)]);
                                         build_recipes_and_cells(r.nm, r.outer, r.cells);
                                         return;
                                     }
                                     r.img = build_host_image(r.nm);
                                     build_update_fields_lists(m.host, r.nm, r.copy, r.coarsen, r.refine);
                                 });
                }
                else
                {
                    for (size_t k : work)
                    {
                        const Mesh& m = *inst_[size_t(act[k])].u->mesh;
                        hip_check(hipMemcpyAsync(host_tags_ + off[k], tag_arena_ + off[k], size_t(m.host.nref), hipMemcpyDeviceToHost, s),
                                  "hipMemcpyAsync(tags)");
                    }
                    ctx_->sync();
                    tt.reset();
                    parallel_for(int(work.size()),
                                 [&](int q)
                                 {
                                     const size_t k      = work[size_t(q)];
                                     const Mesh& m       = *inst_[size_t(act[k])].u->mesh;
                                     const MeshConfig& c = m.host.cfg;
                                     Rebuild& r          = rb[k];
                                     CellArray ca        = update_cell_array_from_tag(m.host, host_tags_ + off[k]);
                                     make_graduation(ca, c.dim, c.grad_width);
                                     if (ca_equal(m.host.ids[CELLS], ca))
                                     {
                                         r.converged = true;
                                         return;
                                     }
                                     r.nm           = build_mesh(c, std::move(ca));
                                     r.nm.graduated = true;
                                     r.img          = build_host_image(r.nm);
                                     build_update_fields_lists(m.host, r.nm, r.copy, r.coarsen, r.refine);
                                 });
                }
            }
            std::vector<size_t> changed;
            for (size_t k = 0; k < na; ++k)
            {
                if (!rb[k].converged)
                {
                    changed.push_back(k);
                }
            }
            if (!changed.empty())
            {
                const size_t nc = changed.size();
                std::vector<std::unique_ptr<DevMesh>> ndev(nc);
                std::vector<std::array<DevWorkList, 3>> ufl(nc);
                {
                    ScopedTimer t(ctx_, "device upload");
                    // each new image (with the update_fields lists, which share its
                    // allocation) is packed on the host in parallel, staged in
                    // pinned memory, then allocated and copied in stream order
                    std::vector<std::unique_ptr<Packer>> packers(nc);
                    std::optional<ScopedTimer> tp(std::in_place, ctx_, "image packing");
                    parallel_for(int(nc),
                                 [&](int q)
                                 {
                                     const size_t j      = size_t(q);
                                     Rebuild& r          = rb[changed[j]];
                                     ndev[j]             = std::make_unique<DevMesh>();
                                     ndev[j]->table_blob = std::make_unique<Blob>();
                                     packers[j]          = std::make_unique<Packer>(*ndev[j]->table_blob);
                                     if (r.device_lists)
                                     {
                                         pack_device_lists(*packers[j], *ndev[j], r.nm, std::move(r.outer), std::move(r.cells), r.dl.data(),
                                                           batch_list_mem);
                                         for (int k = 0; k < 3; ++k)
                                         {
                                             pack_device_list(*packers[j], ufl[j][size_t(k)], dev::DL_UF_COPY + k, r.nm.cfg,
                                                              r.dl[size_t(dev::DL_UF_COPY + k)]);
                                         }
                                         return;
                                     }
                                     pack_image(*packers[j], *ndev[j], r.nm, std::move(r.img));
                                     ufl[j][0].host = std::move(r.copy);
                                     ufl[j][1].host = std::move(r.coarsen);
                                     ufl[j][2].host = std::move(r.refine);
                                     for (DevWorkList& l : ufl[j])
                                     {
                                         packers[j]->list(l);
                                     }
                                 });
                    tp.reset();
                    std::vector<size_t> soff(nc + 1, 0);
                    for (size_t j = 0; j < nc; ++j)
                    {
                        soff[j + 1] = soff[j] + ((ndev[j]->table_blob->host_bytes() + 255) & ~size_t(255));
                    }
                    char* stage = staging_.acquire(soff[nc]);
                    parallel_for(int(nc), [&](int q) { ndev[size_t(q)]->table_blob->stage(stage + soff[size_t(q)]); });
                    for (size_t j = 0; j < nc; ++j)
                    {
                        packers[j]->finish(s);
                    }
                    staging_.release(s);
                    // resolve every new list; nothing runs on them unless all pairs resolve
                    resolve_prog_.clear();
                    std::vector<Chain> chains(nc);
                    for (size_t j = 0; j < nc; ++j)
                    {
                        const Mesh& m       = *inst_[size_t(act[changed[j]])].u->mesh;
                        const MeshConfig& c = m.host.cfg;
                        const dev::Table same[2] = {ndev[j]->table, ndev[j]->table};
                        const dev::Table both[2] = {ndev[j]->table, m.dev->table};
                        const size_t hs          = resolve_prog_.add_aux(same, sizeof(same));
                        const size_t hb          = resolve_prog_.add_aux(both, sizeof(both));
                        Stage st;
                        auto add = [&](DevWorkList& l, size_t h)
                        {
                            dev::Segment g = list_seg(dev::SEG_RESOLVE, c, l, 0, dev::resolve_range(l.d));
                            g.tabs         = reinterpret_cast<const dev::Table*>(uintptr_t(h));
                            g.err          = ctx_->d_err;
                            g.aux          = AUX_TABS;
                            if (g.c1 > g.c0)
                            {
                                st.push_back(g);
                            }
                        };
                        for (DevWorkList* l : ndev[j]->lists())
                        {
                            add(*l, hs);
                        }
                        for (DevWorkList& l : ufl[j])
                        {
                            add(l, hb);
                        }
                        chains[j].push_back(std::move(st));
                        // then the row bases of the cells list's column walks
                        DevWorkList& cl = ndev[j]->cells;
                        if (cl.d.nwalks > 0)
                        {
                            dev::Segment g = list_seg(dev::SEG_WALK_ROWS, c, cl, 0, cl.d.nwalks);
                            chains[j].push_back(Stage{g});
                        }
                    }
                    resolve_prog_.append(chains);
                    resolve_prog_.upload_for(pass_tuning());
                    resolve_prog_.run(pass_tuning());
                    ctx_->check_resolve();
                }
                // update_fields (update_fields.hpp:27-127) into new storage; unp1 is resized (zero-filled)
                {
                    ScopedTimer t(ctx_, "fields update");
                    uf_prog_.clear();
                    std::vector<Chain> chains(nc);
                    std::vector<double*> nu(nc), nv(nc);
                    {
                        ScopedTimer ta(ctx_, "field storage");
                        for (size_t j = 0; j < nc; ++j)
                        {
                            const int64_t n = rb[changed[j]].nm.nref;
                            nu[j]           = static_cast<double*>(dev_alloc(size_t(std::max<int64_t>(n, 1)) * 8, s));
                            nv[j]           = static_cast<double*>(dev_alloc(size_t(std::max<int64_t>(n, 1)) * 8, s));
                        }
                    }
                    std::optional<ScopedTimer> tp(std::in_place, ctx_, "update_fields program");
                    for (size_t j = 0; j < nc; ++j)
                    {
                        const BatchInstance& b = inst_[size_t(act[changed[j]])];
                        const MeshConfig& c    = b.u->mesh->host.cfg;
                        const int64_t n        = rb[changed[j]].nm.nref;
                        Chain ch(1);
                        dev::Segment z0 = make_seg(dev::SEG_FILL_F64, c);
                        z0.out          = nu[j];
                        z0.c1           = n;
                        dev::Segment z1 = z0;
                        z1.out          = nv[j];
                        ch[0]           = {z0, z1};
                        const int ops[3] = {dev::SEG_COPY, dev::SEG_PROJECT_TO, dev::SEG_PREDICT_CHILDREN};
                        for (int q = 0; q < 3; ++q)
                        {
                            dev::Segment g = list_seg(ops[q], c, ufl[j][size_t(q)], 0, ufl[j][size_t(q)].d.ncells);
                            g.out          = nu[j];
                            g.u            = b.u->d;
                            push_stage(ch, g);
                        }
                        chains[j] = std::move(ch);
                    }
                    uf_prog_.append(chains);
                    uf_prog_.upload_for(pass_tuning());
                    uf_prog_.run(pass_tuning());
                    tp.reset();
                    ScopedTimer tr(ctx_, "old storage release");
                    struct Replaced
                    {
                        std::unique_ptr<DevMesh> dev;
                        MRMeshData host;
                        std::array<DevWorkList, 3> uf;
                        Rebuild rb;
                    };
                    std::vector<std::unique_ptr<DevMesh>> old_dev(nc);
                    std::vector<MRMeshData> old_host(nc);
                    for (size_t j = 0; j < nc; ++j)
                    {
                        BatchInstance& b = inst_[size_t(act[changed[j]])];
                        Mesh& m          = *b.u->mesh;
                        const int64_t n  = rb[changed[j]].nm.nref;
                        // old storage and old image are released in stream order after the program
                        dev_free(b.u->d, s);
                        dev_free(b.unp1->d, s);
                        old_host[j] = std::move(m.host);
                        old_dev[j]  = std::move(m.dev);
                        m.replace(std::move(rb[changed[j]].nm), std::move(ndev[j]));
                        b.u->d          = nu[j];
                        b.unp1->d       = nv[j];
                        b.u->size       = n;
                        b.unp1->size    = n;
                        b.u->version    = m.version;
                        b.unp1->version = m.version;
                    }
                    // the old versions and the update_fields lists are freed
                    // (in parallel) while the next launches run: their device
                    // blocks go back in stream order, after the work queued now
                    for (size_t j = 0; j < nc; ++j)
                    {
                        auto* r = new Replaced{std::move(old_dev[j]), std::move(old_host[j]), std::move(ufl[j]), std::move(rb[changed[j]])};
                        replaced_.emplace_back(r, [](void* p) { delete static_cast<Replaced*>(p); });
                    }
                }
            }
            std::vector<int> next;
            for (size_t k : changed)
            {
                next.push_back(act[k]);
            }
            act.swap(next);
        }
        return calls;
    }

    ProgramBuf& Batch::fv_program()
    {
        // identity of the program: image serials, storages and boundary values
        std::vector<const void*> key;
        key.reserve(inst_.size() * 4);
        for (const BatchInstance& b : inst_)
        {
            key.push_back(reinterpret_cast<const void*>(uintptr_t(b.u->mesh->dev->serial)));
            key.push_back(b.u->d);
            key.push_back(b.unp1->d);
            key.push_back(reinterpret_cast<const void*>(uintptr_t(b.u->has_bc ? 1 : 0)));
            uint64_t bits;
            std::memcpy(&bits, &b.u->bc_value, 8);
            key.push_back(reinterpret_cast<const void*>(uintptr_t(bits)));
        }
        for (int q = 0; q < 2; ++q)
        {
            if (fv_key_[q] == key)
            {
                return *fv_prog_[q];
            }
        }
        const int q   = fv_next_;
        fv_next_      = 1 - fv_next_;
        ProgramBuf& p = *fv_prog_[q];
        p.clear();
        std::vector<Chain> chains(inst_.size());
        for (size_t i = 0; i < inst_.size(); ++i)
        {
            BatchInstance& b    = inst_[i];
            DevMesh& dm         = *b.u->mesh->dev;
            const MeshConfig& c = b.u->mesh->host.cfg;
            Chain ch;
            ghost_chain(ch, dm, c, b.u->d, b.u->has_bc, b.u->bc_value);
            dev::Segment g = list_seg(dev::SEG_UPWIND, c, dm.cells, 0, dm.cells.d.ncells);
            g.u            = b.u->d;
            g.out          = b.unp1->d;
            g.a0           = b.a[0];
            g.a1           = b.a[1];
            g.a2           = c.dim == 3 ? b.a[2] : 0.0;
            g.dt           = b.dt;
            push_stage(ch, g);
            chains[i] = std::move(ch);
        }
        p.append(chains);
        p.upload();
        fv_key_[q] = std::move(key);
        return p;
    }

    Tuning Batch::pass_tuning() const
    {
        // the programs of an adaptation pass run once: one cell per thread
        // (no wave items to build on the host), measured 4.5 ms faster per
        // step of 256 instances than wave items; SMR_PASS_CPT sets another
        // cells-per-thread count (0: wave items)
        static const int cpt = []
        {
            const char* e = std::getenv("SMR_PASS_CPT");
            return e ? std::max(0, std::min(std::atoi(e), 64)) : 1;
        }();
        Tuning t = tuning_;
        t.cpt    = cpt;
        return t;
    }

    void Batch::release_replaced()
    {
        if (replaced_.empty())
        {
            return;
        }
        ScopedTimer t(nullptr, "replaced versions freed");
        parallel_for(int(replaced_.size()), [&](int i) { replaced_[size_t(i)].reset(); });
        replaced_.clear();
    }

    void Batch::advance()
    {
        check_instances();
        std::vector<int> all(inst_.size());
        for (size_t i = 0; i < all.size(); ++i)
        {
            all[i] = int(i);
        }
        ensure_images(all);
        {
            ScopedTimer t(ctx_, "ghost update + field expressions");
            fv_program().run(tuning_);
            // the versions the adaptation replaced, freed on the host while the
            // advance program runs
            release_replaced();
        }
        for (BatchInstance& b : inst_)
        {
            std::swap(b.u->d, b.unp1->d);
            std::swap(b.u->size, b.unp1->size);
            std::swap(b.u->version, b.unp1->version);
        }
    }

    int Batch::step(double eps, double regularity)
    {
        const int calls = adapt(eps, regularity);
        advance();
        return calls;
    }

    // ------------------------------------------------------------------
    // autotuner
    // ------------------------------------------------------------------
    bool Batch::load_tuning(const std::string& cache_file, Tuning& t)
    {
        if (cache_file.empty())
        {
            return false;
        }
        int dev = 0;
        (void)hipGetDevice(&dev);
        const std::string arch = gpu_arch(dev);
        std::ifstream in(cache_file);
        std::string line;
        while (std::getline(in, line))
        {
            std::istringstream ls(line);
            std::string a;
            Tuning v;
            if (ls >> a >> v.block >> v.cpt >> v.mode && a == arch)
            {
                t = v;
                return true;
            }
        }
        return false;
    }

    Tuning Batch::autotune(int reps, const std::string& cache_file)
    {
        check_instances();
        std::vector<int> all(inst_.size());
        for (size_t i = 0; i < all.size(); ++i)
        {
            all[i] = int(i);
        }
        ensure_images(all);
        ProgramBuf& p = fv_program();
        Tuning best   = tuning_;
        double best_s = -1;
        // Each repetition uploads the program again before running it, as a step
        // whose meshes changed does: a replayed graph must then be captured and
        // instantiated anew, which the wall time of the repetition includes.
        for (int mode : {MODE_FUSED, MODE_STAGES, MODE_GRAPH})
        {
            for (int block : {64, 128, 256, 512, 1024})
            {
                if (mode == MODE_FUSED && max_blocks_cached(block) <= 0)
                {
                    continue;
                }
                for (int cpt : {0, 1, 2, 4, 8})
                {
                    // wave items (cpt 0) are launched in blocks of 256 whatever the block
                    // size, and the fused launch walks cells
                    if (cpt == 0 && (block != 256 || mode == MODE_FUSED))
                    {
                        continue;
                    }
                    const Tuning t{block, cpt, mode};
                    p.upload();
                    p.run(t); // warm-up
                    ctx_->sync();
                    const auto t0 = std::chrono::steady_clock::now();
                    for (int r = 0; r < reps; ++r)
                    {
                        p.upload();
                        p.run(t);
                    }
                    ctx_->sync();
                    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                    if (best_s < 0 || sec < best_s)
                    {
                        best_s = sec;
                        best   = t;
                    }
                }
            }
        }
        tuning_ = best;
        if (!cache_file.empty())
        {
            const std::string arch = gpu_arch(ctx_->device);
            std::vector<std::string> keep;
            {
                std::ifstream in(cache_file);
                std::string line;
                while (std::getline(in, line))
                {
                    std::istringstream ls(line);
                    std::string a;
                    if (ls >> a && a != arch)
                    {
                        keep.push_back(line);
                    }
                }
            }
            std::ofstream out(cache_file, std::ios::trunc);
            for (const std::string& l : keep)
            {
                out << l << "\n";
            }
            out << arch << " " << best.block << " " << best.cpt << " " << best.mode << "\n";
        }
        return best;
    }
} // namespace smr
