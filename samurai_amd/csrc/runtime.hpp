// Runtime of the MI355X MR step: device images of the meshes, fields, and
// the host/device orchestration of update_ghost_mr, the Harten adaptation,
// update_fields and the upwind step.
#pragma once

#include "device/device_api.hpp"
#include "host/convection.hpp"
#include "host/programs.hpp"

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace smr
{
    struct Ctx;

    void hip_check(hipError_t e, const char* what);

    // One device allocation holding several host arrays (uploaded together).
    // The arrays are laid out when added and copied once, when the blob is
    // staged (into pinned memory) or copied from a host buffer of its own.
    class Blob
    {
      public:
        Blob() = default;
        ~Blob();
        Blob(const Blob&)            = delete;
        Blob& operator=(const Blob&) = delete;

        // reserves space for n bytes of src (256-byte aligned) and returns the
        // offset; src must stay valid and unchanged until stage() or copy()
        size_t add(const void* src, size_t n);
        // the same for a vector the blob keeps until it is staged or copied
        template <class T>
        size_t add(std::vector<T>&& v)
        {
            auto kept        = std::make_shared<std::vector<T>>(std::move(v));
            const size_t off = add(kept->data(), kept->size() * sizeof(T));
            keep_.push_back(std::move(kept));
            return off;
        }
        // n bytes of device-only memory (no host copy, not uploaded, contents
        // undefined until a kernel writes them); returns an offset for sptr()
        size_t scratch(size_t n);
        // stream-ordered allocation + copy; the host staging copy is kept until
        // the blob dies, so no host synchronization is needed here
        void upload(hipStream_t s);
        // upload in two steps, to bind device pointers between them
        void alloc(hipStream_t s);
        void copy();
        // writes the host arrays into dst (pinned memory of host_bytes() bytes
        // that stays untouched until the copy has run); copy() then reads dst
        void stage(char* dst);
        char* dptr(size_t off) const { return d_ + off; }
        char* sptr(size_t off) const { return d_ + scratch_base() + off; }
        size_t host_bytes() const { return host_size_; }
        size_t bytes() const { return scratch_base() + scratch_; }

      private:
        struct Part
        {
            const void* src;
            size_t n, off;
        };
        size_t scratch_base() const { return (host_size_ + 255) & ~size_t(255); }
        void write_parts(char* dst);
        std::vector<Part> parts_;
        std::vector<std::shared_ptr<void>> keep_;
        std::unique_ptr<char[]> host_; // when copied without staging
        size_t host_size_    = 0;
        const char* staged_  = nullptr;
        size_t scratch_      = 0;
        char* d_           = nullptr;
        hipStream_t stream_ = nullptr;
    };

    // Pinned host staging reused across uploads: acquire() returns at least n
    // bytes once the copies issued from its previous contents have run;
    // release() marks the copies just issued on s
    class PinnedStaging
    {
      public:
        PinnedStaging() = default;
        ~PinnedStaging();
        PinnedStaging(const PinnedStaging&)            = delete;
        PinnedStaging& operator=(const PinnedStaging&) = delete;
        char* acquire(size_t n);
        void release(hipStream_t s);

      private:
        char* h_          = nullptr;
        size_t cap_       = 0;
        hipEvent_t done_  = nullptr;
        bool pending_     = false;
    };

    // Device memory used in the order of stream s.  A freed block is kept in a
    // cache of s and handed to a later dev_alloc on s: every earlier use of it
    // was queued on s before every later one, so no synchronization is needed
    // (a free costs ~30 us in the HIP runtime, a cache hit nothing).  Misses
    // come from the device's stream-ordered pool; the cache of a stream is
    // trimmed above kDevCacheLimit bytes and released by dev_cache_release.
    void* dev_alloc(size_t bytes, hipStream_t s);
    void dev_free(void* p, hipStream_t s);
    void dev_cache_release(hipStream_t s);
    constexpr size_t kDevCacheLimit = size_t(4) << 30;

    // Device memory from dev_alloc, released in the order of its stream.
    struct DevBuffer
    {
        void* p         = nullptr;
        hipStream_t s   = nullptr;
        DevBuffer()     = default;
        DevBuffer(const DevBuffer&)            = delete;
        DevBuffer& operator=(const DevBuffer&) = delete;
        DevBuffer(DevBuffer&& o) noexcept : p(o.p), s(o.s) { o.p = nullptr; }
        DevBuffer& operator=(DevBuffer&& o) noexcept
        {
            if (this != &o)
            {
                reset();
                p   = o.p;
                s   = o.s;
                o.p = nullptr;
            }
            return *this;
        }
        ~DevBuffer() { reset(); }
        void reset()
        {
            if (p)
            {
                dev_free(p, s);
                p = nullptr;
            }
        }
        void* get(size_t bytes, hipStream_t stream)
        {
            reset();
            s = stream;
            p = dev_alloc(bytes, stream);
            return p;
        }
    };

    struct DevWorkList
    {
        WorkList host;
        dev::List d{};
        std::vector<int32_t> hint;
        std::unique_ptr<Blob> blob;
        // built on the device: host holds the taps and the first run of every
        // level only, level_cell the first cell of every level (MAXL + 1)
        bool device_built = false;
        std::vector<int64_t> level_cell;
        // cell range [c0, c1) of the runs of levels [l0, l1)
        void cells_of_levels(int l0, int l1, int64_t& c0, int64_t& c1) const;
    };

    struct DevRecipes
    {
        Recipes host;
        std::unique_ptr<Blob> blob;
        const int64_t* dst = nullptr;
        const int32_t* op  = nullptr;
        const int64_t* sp  = nullptr;
        const int64_t* src = nullptr;
    };

    // A chain of dependent stages launched as: the large stages one launch
    // each, every run of consecutive small stages (at most kSeqSmall cells
    // each) one single-block launch of dev::sequence.
    struct SeqPlan
    {
        struct Step
        {
            bool seq;
            int i0, i1;
        };
        std::vector<dev::SeqItem> items; // kept for the asynchronous upload
        std::vector<Step> steps;
        DevBuffer d_items;
        bool built = false;
        void build(std::vector<dev::SeqItem>&& all, hipStream_t s);
    };
    constexpr int64_t kSeqSmall = 2048;

    // Device image of one mesh version.
    struct DevMesh
    {
        DevMesh();
        uint64_t serial; // unique per image of the process
        std::unique_ptr<Blob> table_blob;
        dev::Table table{};
        DevWorkList cells;
        bool has_ghost = false, has_adapt = false;
        std::vector<DevRecipes> outer;
        DevWorkList projection, prediction;
        DevWorkList detail, criteria, maximum;
        // flux-based convection terms of the last velocity used on this image
        struct Conv
        {
            double v[3] = {0, 0, 0};
            std::unique_ptr<Blob> blob;
            const int64_t *row = nullptr, *ptr = nullptr, *src = nullptr;
            const double* coef = nullptr;
            const int8_t* dir  = nullptr;
            int64_t nrows      = 0;
        };
        std::unique_ptr<Conv> conv;
        // the non-linear (Burgers) convection terms of this image
        std::unique_ptr<Conv> burgers;
        // the heterogeneous (velocity field) convection terms of this image
        std::unique_ptr<Conv> heterogeneous;
        // the WENO5 convection terms of this image (one-level meshes of ghost width 3)
        std::unique_ptr<Conv> weno5;
        // launch plans of update_ghost_mr and of the tag finalization
        SeqPlan ghost_plan, max_plan;
        SeqPlan ghost_plan_nobc; // update_ghost_mr of a field without a B.C.
        // the leaves as per-level bitmaps (device mesh update), built on first use
        DevBuffer leaf_bits;
        bool has_leaf_bits = false;
        // update_fields lists (copy, coarsen, refine) from the mesh this image
        // replaced, packed with the image when it was built for a mesh change
        bool has_uf = false;
        DevWorkList uf[3];
        // the arrays of the work lists built on the device, and the padded
        // bitmaps of the mesh ids of this version (the old mesh of the next
        // device-built update_fields lists)
        std::shared_ptr<DevBuffer> list_mem; // shared by the meshes of one batched update
        DevBuffer mesh_sets;
        bool has_mesh_sets = false;
        // the work lists of a complete image (cells first)
        std::vector<DevWorkList*> lists() { return {&cells, &projection, &prediction, &detail, &criteria, &maximum}; }
    };

    // Packs tables, work lists and recipes into one Blob; the device pointers
    // are bound when finish() has allocated it (then the data is copied).
    class Packer
    {
      public:
        explicit Packer(Blob& b) : b_(b) {}
        void table(const MRMeshData& m, dev::Table& t);
        void list(DevWorkList& l);
        // several lists, their host arrays prepared in parallel
        void lists(const std::vector<DevWorkList*>& ls);
        void recipes(DevRecipes& r);
        // a list whose runs, prefixes and hints the device built (at iv,
        // pref, hint); its taps come from l.host
        void device_list(DevWorkList& l, const int32_t* iv, const int64_t* pref, const int32_t* hint, int64_t nruns, int64_t ncells);
        void finish(hipStream_t s);

      private:
        struct ListPrep
        {
            std::vector<int8_t> taps;
            std::vector<int32_t> tiles, next, walks;
            int64_t nrows_total = 0;
            int32_t walk_stride = 0;
        };
        static ListPrep prepare_list(DevWorkList& l);
        void commit_list(DevWorkList& l, ListPrep&& prep);
        Blob& b_;
        std::vector<std::function<void()>> binds_;
    };

    struct Mesh
    {
        Ctx* ctx = nullptr;
        MRMeshData host;
        uint64_t version = 0;
        std::unique_ptr<DevMesh> dev;

        DevMesh& device();          // table + cells list (uploaded and resolved)
        DevMesh& device_ghost();    // + ghost program
        DevMesh& device_adapt();    // + adapt program
        // the host arrays of work lists the device built (for host code that
        // reads the lists' runs)
        void ensure_host_lists();
        void replace(MRMeshData&& m, std::unique_ptr<DevMesh>&& d);
    };

    struct Field
    {
        Mesh* mesh = nullptr;
        std::string name;
        double* d       = nullptr;
        int64_t size    = 0;
        uint64_t version = 0;
        bool has_bc     = false;
        double bc_value = 0;
        ~Field();
        void allocate(bool zero);
        void check_current() const;
    };

    struct AdaptState
    {
        Field* field   = nullptr;
        // every adapted field, `field` first (make_MRAdapt(u, v...),
        // mr/adapt.hpp:391-397): their details join the criteria and they are
        // transferred to every new mesh
        std::vector<Field*> fields;
        int ncomp() const { return fields.empty() ? 1 : int(fields.size()); }
        double* detail = nullptr;
        uint8_t* tag   = nullptr;
        int64_t capacity = 0;
        // pinned copy of the tags of the last harten call (on the mesh they
        // were computed on); the details of that call stay in `detail`
        uint8_t* host_tag         = nullptr;
        int64_t host_tag_capacity = 0;
        int64_t last_size         = 0;
        bool host_tag_valid       = false; // else the tags are only in `tag`
        // mra_config().relative_detail(true) (mr/config.hpp:36-45): the
        // details divided by max |field| over the leaves before the criteria
        bool relative_detail = false;
        // the --refine-boundary option (arguments.hpp; mr/adapt.hpp:341-346)
        bool refine_boundary = false;
        unsigned long long* rel_max = nullptr; // one device word per field
        int rel_capacity            = 0;
        ~AdaptState();
        void copy_last(uint8_t* tags, double* details, hipStream_t s) const;
    };

    struct Ctx
    {
        int device         = -1;
        hipStream_t stream = nullptr;
        int* d_err         = nullptr; // two ints: resolve error, then a scratch flag
        // pinned host memory the mesh images are staged in before their copy
        std::unique_ptr<PinnedStaging> staging = std::make_unique<PinnedStaging>();
        // pinned host memory the runs of a new mesh are read back into
        std::unique_ptr<PinnedStaging> readback = std::make_unique<PinnedStaging>();
        // recorded after a new mesh's runs are read back, so that the host can
        // complete the mesh while later kernels run
        hipEvent_t runs_back = nullptr;
        void mark_runs_back();
        ~Ctx();
        int* d_flag() const { return d_err + 1; }
        bool has_device() const { return device >= 0; }
        void require_device() const;
        void sync();
        void check_resolve();
    };

    // Phase timers named after the reference's ScopedTimer names (timers.hpp,
    // mr/adapt.hpp:151-384, update_ghost_mr.hpp:263, field_base.hpp:234).  When
    // enabled, each phase synchronizes the stream on entry and exit so that its
    // wall time includes the device work it launched.
    struct Timers
    {
        bool enabled = false;
        std::vector<std::pair<std::string, std::pair<double, int64_t>>> entries;
        void add(const std::string& name, double seconds);
        std::string report() const;
    };
    Timers& timers();

    class ScopedTimer
    {
      public:
        ScopedTimer(Ctx* ctx, const char* name);
        ~ScopedTimer();

      private:
        Ctx* ctx_;
        const char* name_;
        double t0_ = 0;
    };

    std::unique_ptr<DevMesh> make_dev_mesh(Ctx& ctx, const MRMeshData& m);

    // host half of a complete device image (thread-safe: no HIP calls)
    struct HostImage
    {
        WorkList cells;
        GhostProgram ghost;
        AdaptProgram adapt;
        // update_fields (copy, coarsen, refine) from an older mesh, when given
        bool has_uf = false;
        WorkList uf[3];
    };
    // with old: also the update_fields lists from old to m
    HostImage build_host_image(const MRMeshData& m, const MRMeshData* old = nullptr);
    // packs a complete image (table, cells, ghost and adapt programs) into p
    void pack_image(Packer& p, DevMesh& dm, const MRMeshData& m, HostImage&& h);
    // device half: upload and resolve (check = wait for and verify the resolve)
    // old_table: the table of the mesh the image replaces (resolves the
    // update_fields lists of an image built with them)
    std::unique_ptr<DevMesh> upload_image(Ctx& ctx, const MRMeshData& m, HostImage&& h, bool check,
                                          const dev::Table* old_table = nullptr);
    // DevWorkList upload + resolve against the given tables
    void upload_list(Ctx& ctx, DevWorkList& l, const dev::Table& t0, const dev::Table& t1);

    struct DeviceLists;

    // bitmap layout of the levels of a mesh configuration
    dev::Bits make_bits(const MeshConfig& cfg);
    // update_cell_array_from_tag + make_graduation on the device from tags
    // indexed by dm's storage (graduation.hpp:573-843): false if the leaves
    // do not change, else true with the new leaves in out
    // With mesh: the new mesh's ids are built on the device too
    // (update_sub_mesh_impl, mr/mesh.hpp:221-455) and *mesh is the complete
    // new mesh (out is then left empty); SMR_HOST_SUBMESH=1 builds them on the
    // host from out instead.
    // With lists too: the work lists of the new mesh are built on the device
    // (with the update_fields lists from dm's mesh) into *lists.
    bool device_mesh_update(Ctx& ctx, DevMesh& dm, const MeshConfig& cfg, const uint8_t* tag, CellArray& out, const MRMeshData* current = nullptr,
                            MRMeshData* mesh = nullptr, DeviceLists* lists = nullptr);
    // SMR_HOST_SUBMESH=1: the mesh ids are built on the host from the leaves
    bool host_submesh();
    // SMR_HOST_LISTS=1: the work lists of a mesh change are built on the host
    bool host_lists();

    // the work lists of a new mesh built on the device with its ids: the
    // arrays of every list (runs, prefixes, hints) in mem, per list its counts
    // and the first run and cell of every level; sets: the padded ids
    struct DeviceLists
    {
        struct One
        {
            const int32_t* iv   = nullptr;
            const int64_t* pref = nullptr;
            const int32_t* hint = nullptr;
            int64_t nruns = 0, ncells = 0;
            std::vector<int64_t> level_run, level_cell;
        };
        int nlists = 0;
        One l[dev::DL_N];
        std::shared_ptr<DevBuffer> mem;
        DevBuffer sets;
    };
    // a device-built list from its row of a list table (packed runs << 32 |
    // cells scans at the first word of every level): its counts and
    // per-level offsets into o, its first run and cell and counts into w;
    // returns the bytes its arrays take; throws if the row is not increasing
    size_t list_from_table(const uint64_t* t, DeviceLists::One& o, dev::ListOut& w);
    // points o and w at the arrays of that list laid out from d
    void bind_list_arrays(char* d, DeviceLists::One& o, dev::ListOut& w);
    // packs the image of a mesh whose work lists the device built: the table,
    // the recipes and the cells list from the host, the five program lists
    // l[DL_PROJECTION .. DL_MAXIMUM] (whose arrays mem holds) as they are
    void pack_device_lists(Packer& p, DevMesh& dm, const MRMeshData& m, std::vector<Recipes>&& outer, WorkList&& cells, const DeviceLists::One* l,
                           std::shared_ptr<DevBuffer> mem);
    // one device-built list of that kind
    void pack_device_list(Packer& p, DevWorkList& w, int kind, const MeshConfig& cfg, const DeviceLists::One& o);
    // the image of a mesh whose projection, prediction, detail, criteria,
    // maximum and update_fields lists the device built (dl); the host gives
    // the recipes and the cells list
    // 0 when the projection, prediction, detail, criteria and maximum lists
    // the device built for m's image equal the host's (throws where they
    // differ), 1 when the image has no device-built lists
    int check_device_lists(Mesh& m);
    std::unique_ptr<DevMesh> upload_image_device_lists(Ctx& ctx, const MRMeshData& m, std::vector<Recipes>&& outer, WorkList&& cells,
                                                       DeviceLists&& dl, const dev::Table* old_table);
    // the meshes whose ids are the runs (x0, x1, y, z; dev::SM_RUN int32
    // each) of a device mesh construction: id `id` of mesh q at level l is
    // the runs [seg[b + l], seg[b + l + 1]), b = (SM_NSETS q + id) SM_SEG;
    // each mesh is completed with its corners and renumbering
    void meshes_from_runs(const std::vector<const MeshConfig*>& cfgs, const int32_t* runs, const int32_t* seg, const std::vector<MRMeshData*>& out);

    // criteria thresholds of harten iteration `ite` (mr/adapt.hpp:325-339)
    dev::Thresholds make_thresholds(const MeshConfig& cfg, int ite, double eps, double regularity);

    void update_ghost_mr(Field& f);
    void upwind(const double* a, double dt, Field& u, Field& unp1);
    // out = u - a * v on the cells (the field expression of the flux-based demos)
    void sub_scaled(Field& u, double a, Field& v, Field& out);
    // README.md:135-154: unp1 = u - dt / dx * (u - u(i-1) + u - u(j-1)) on the cells (2D)
    void upwind_hand_loop(double dt, Field& u, Field& unp1);
    // out = make_convection_upwind(v)(u), after update_ghost_mr(u)
    void convection_upwind(const double* v, Field& u, Field& out);
    // unp1 = u - dt * upwind_scalar_burgers(k, u), 2D, ghosts of u updated beforehand
    void upwind_scalar_burgers(const double* k, double dt, Field& u, Field& unp1);
    // out = make_convection_upwind<Field>()(u) (scalar, non-linear), after update_ghost_mr(u)
    void convection_upwind_nonlinear(Field& u, Field& out);
    // out = make_convection_upwind<Field>(velocity_field)(u): velocity given as
    // dim scalar component fields on u's mesh; their ghosts and u's are updated
    void convection_upwind_heterogeneous(Field* const* velocity, Field& u, Field& out);
    // out = make_convection_weno5(v)(u), after update_ghost_mr(u) (one-level
    // mesh of stencil radius 3)
    void convection_weno5(const double* v, Field& u, Field& out);
    // Adapt::operator() (mr/adapt.hpp:147-195); returns the number of harten calls
    int adapt(AdaptState& st, double eps, double regularity, int max_iterations = -1);
    // one harten call; true at the fixed point
    bool harten(AdaptState& st, int ite, double eps, double regularity);
    // the device half of a harten call (mr/adapt.hpp:276-358): tag init, ghost
    // update, details, criteria and tag finalization of iteration ite, into
    // tag and det (device arrays of the mesh's nref slots); no host copy
    void harten_tags(Field& f, int ite, double eps, double regularity, uint8_t* tag, double* det);
    // harten_tags over nf fields on one mesh (fs[0] first): ghosts of each,
    // details of field k into det + k * nref, criteria over all of them
    // (rel_max: one device word per field for the relative detail, null for
    // the absolute one)
    void harten_tags(Field* const* fs, int nf, int ite, double eps, double regularity, uint8_t* tag, double* det,
                     unsigned long long* rel_max = nullptr, bool refine_boundary = false);
    // harten_tags into a caller's device tag array, details in st's buffer
    void adapt_tags(AdaptState& st, int ite, double eps, double regularity, uint8_t* tag);

    // slots (doubles) of the static device array adapt_tags_noalloc keeps its
    // details in (16 MiB per GPU, held until the process ends)
    constexpr int64_t kNoAllocSlots = int64_t(1) << 21;
    // takes the device address of that array on `device` (current); once per device
    void noalloc_init(int device);
    // harten_tags with the details in the static array: no allocation.  The
    // mesh's ghost and adapt programs must exist (built by an earlier
    // allocating call on this mesh version); nref <= kNoAllocSlots
    void adapt_tags_noalloc(AdaptState& st, int ite, double eps, double regularity, uint8_t* tag);
} // namespace smr
