// engine.cpp - host side of the engine: buffers by id, kernels by name, the
// fused hipGraph fast path, and the C ABI declared in include/fibrenet.h.
//
// Device state is structure-of-arrays (see kernels.h).  The reference's AoS
// records are the wire format, converted on upload and read-back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "engine.h"
#include "fibrenet.h"

namespace {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCK(x)                                                                           \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) throw Error(FN_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

std::string fmt(const char *f, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    return buf;
}

template <typename X>
struct DArr {
    X *p = nullptr;
    size_t n = 0;
    DArr() = default;
    DArr(const DArr &) = delete;
    DArr &operator=(const DArr &) = delete;
    ~DArr() { release(); }
    void alloc(size_t k)
    {
        release();
        if (k) HIPCK(hipMalloc(&p, k * sizeof(X)));
        n = k;
    }
    void ensure(size_t k)
    {
        if (k > n) alloc(k);
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void put(const std::vector<X> &h, hipStream_t s)
    {
        ensure(h.size());
        if (!h.empty()) HIPCK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice, s));
    }
    void get(std::vector<X> &h, size_t k, hipStream_t s) const
    {
        h.resize(k);
        if (k) HIPCK(hipMemcpyAsync(h.data(), p, k * sizeof(X), hipMemcpyDeviceToHost, s));
    }
};

enum KernelId {
    K_ADVANCE_POSITION, K_RESET_GRID, K_RESET_LIST, K_FILL_GRID, K_BUILD_LIST,
    K_RESOLVE_CONTACTS, K_ADVANCE_VELOCITY, K_RESET_ENERGY, K_RESOLVE_SPRINGS, K_RESOLVE_ANGLES, K_COUNT
};

// Kernel names registered by SimulationManager::ReadKernelData (SimulationManager.cpp:1191-1200).
const char *kKernelNames[K_COUNT] = {"AdvancePosition", "ResetGrid", "ResetContactList", "FillGrid",
                                     "BuildContactList", "ResolveContacts", "AdvanceVelocity",
                                     "ResetBeadEnergy", "ResolveSprings", "ResolveAngles"};

// Argument contract of each kernel (kernel.cl signatures, bound at
// SimulationManager.cpp:1236-1304): FN_ARG_* kind, and the buffer record kind.
struct ArgSpec { int kind; int buf; };
const std::vector<ArgSpec> kContract[K_COUNT] = {
    /* AdvancePosition  */ {{FN_ARG_BUFFER, FN_BUF_BEAD}, {FN_ARG_BUFFER, FN_BUF_BEAD_INFO}, {FN_ARG_SETTINGS, 0},
                            {FN_ARG_INT, 0}},
    /* ResetGrid        */ {{FN_ARG_BUFFER, FN_BUF_INT}, {FN_ARG_INT, 0}},
    /* ResetContactList */ {{FN_ARG_BUFFER, FN_BUF_INT}, {FN_ARG_INT, 0}},
    /* FillGrid         */ {{FN_ARG_BUFFER, FN_BUF_BEAD}, {FN_ARG_BUFFER, FN_BUF_INT}, {FN_ARG_BUFFER, FN_BUF_GRID_CELL},
                            {FN_ARG_SETTINGS, 0}, {FN_ARG_INT, 0}},
    /* BuildContactList */ {{FN_ARG_BUFFER, FN_BUF_INT}, {FN_ARG_BUFFER, FN_BUF_GRID_CELL},
                            {FN_ARG_BUFFER, FN_BUF_GRID_CELL_NH}, {FN_ARG_BUFFER, FN_BUF_INT},
                            {FN_ARG_BUFFER, FN_BUF_PARTICLE_NH}, {FN_ARG_BUFFER, FN_BUF_BEAD},
                            {FN_ARG_BUFFER, FN_BUF_BEAD_INFO}, {FN_ARG_BUFFER, FN_BUF_PARTICLE_NH},
                            {FN_ARG_BUFFER, FN_BUF_PAIR}, {FN_ARG_SETTINGS, 0}, {FN_ARG_INT, 0}},
    /* ResolveContacts  */ {{FN_ARG_BUFFER, FN_BUF_INT}, {FN_ARG_BUFFER, FN_BUF_PARTICLE_NH}, {FN_ARG_BUFFER, FN_BUF_BEAD},
                            {FN_ARG_BUFFER, FN_BUF_BEAD_INFO}, {FN_ARG_BUFFER, FN_BUF_PARTICLE_NH},
                            {FN_ARG_BUFFER, FN_BUF_PAIR}, {FN_ARG_SETTINGS, 0}, {FN_ARG_INT, 0}},
    /* AdvanceVelocity  */ {{FN_ARG_BUFFER, FN_BUF_BEAD}, {FN_ARG_BUFFER, FN_BUF_BEAD_INFO}, {FN_ARG_INT, 0}, {FN_ARG_INT, 0}},
    /* ResetBeadEnergy  */ {{FN_ARG_BUFFER, FN_BUF_BEAD}, {FN_ARG_INT, 0}},
    /* ResolveSprings   */ {{FN_ARG_BUFFER, FN_BUF_BEAD}, {FN_ARG_BUFFER, FN_BUF_SPRING}, {FN_ARG_INT, 0}},
    /* ResolveAngles    */ {{FN_ARG_BUFFER, FN_BUF_ANGLE}, {FN_ARG_BUFFER, FN_BUF_SPRING}, {FN_ARG_BUFFER, FN_BUF_BEAD},
                            {FN_ARG_INT, 0}},
};

// Index of the argument holding the work-item count / settings of each kernel.
const int kCountArg[K_COUNT] = {3, 1, 1, 4, 10, 7, 2, 1, 2, 3};
const int kSettingsArg[K_COUNT] = {2, -1, -1, 3, 9, 6, -1, -1, -1, -1};

int record_size(int kind)
{
    switch (kind) {
    case FN_BUF_BEAD: return sizeof(fn_bead);
    case FN_BUF_BEAD_INFO: return sizeof(fn_bead_info);
    case FN_BUF_SPRING: return sizeof(fn_spring);
    case FN_BUF_ANGLE: return sizeof(fn_angle);
    case FN_BUF_PAIR: return sizeof(fn_pair);
    case FN_BUF_INT: return sizeof(int);
    case FN_BUF_GRID_CELL: return sizeof(fn_grid_cell);
    case FN_BUF_GRID_CELL_NH: return sizeof(fn_grid_cell_nh);
    case FN_BUF_PARTICLE_NH: return sizeof(fn_particle_nh);
    default: return -1;
    }
}

// OpenclManager.h:125 pads every buffer to a multiple of the work-group size (64).
int pad64(int n) { return ((n + 63) / 64) * 64; }

struct Buf { int kind, esize, elements, length; };

struct Arg {
    int kind = 0;
    int ival = 0;
    fn_settings s{};
};

struct KernelRec {
    bool registered = false;
    int range = 0;
    std::map<int, Arg> args;
};

fnk::Box make_box(const fn_settings &s)
{
    fnk::Box b{};
    b.bc = s.bc_type;
    b.cx = s.cells_x;
    b.cy = s.cells_y;
    b.cz = s.cells_z;
    b.cxy = s.cells_xy;
    b.types = s.types_num;
    b.lx = s.box_size.x;
    b.ly = s.box_size.y;
    b.lz = s.box_size.z;
    b.sx = s.grid_step.x;
    b.sy = s.grid_step.y;
    b.sz = s.grid_step.z;
    b.wall_k = s.wall_contact_stiffness;
    b.hbs = s.half_box_sq;
    b.cutoff_sq = s.cutoff_sq;
    b.damping = s.damping;
    b.gx = s.gravity.x;
    b.gy = s.gravity.y;
    b.gz = s.gravity.z;
    b.dt = s.dt;
    return b;
}

}  // namespace

// ----------------------------------------------------------------------------------------
struct fn_engine {
    virtual ~fn_engine() = default;
    std::string err;
    fn_log_fn log_fn = nullptr;
    void *log_user = nullptr;
    int device = 0;
    int precision = 0;

    void log(const std::string &m)
    {
        if (log_fn) log_fn(m.c_str(), log_user);
        else fprintf(stderr, "fibrenet: %s\n", m.c_str());
    }

    virtual int create_buffer(int kind, int esize, int n, const void *host) = 0;
    virtual void write_buffer(int id, const void *host, int n_padded, int esize) = 0;
    virtual void read_buffer(int id, void *host) = 0;
    virtual const Buf *buffer(int id) const = 0;
    virtual void reg(const char *name) = 0;
    virtual void set_arg(const char *kernel, int idx, int kind, const void *value, int size) = 0;
    virtual void set_range(const char *kernel, int range) = 0;
    virtual void launch(const char *kernel) = 0;
    virtual void finish() = 0;
    virtual void run_steps(int n, int guf, bool spr, bool ang) = 0;
    virtual void export_cells(int *out) = 0;
    virtual long long export_pairs(int *out, long long cap) = 0;
    virtual void export_breaks(int *out) = 0;
    virtual void stats(fn_stats *out) = 0;
    virtual void set_timing(bool on) = 0;
    virtual void set_option(const std::string &name, int value) = 0;
    virtual int get_option(const std::string &name) = 0;
    virtual void timing(fn_timing *out) = 0;
    virtual void copy_probe(long long bytes, double *ms) = 0;
    virtual void load_ensemble(const fn_network *nets, int n) = 0;
    virtual void reload_ensemble(const fn_network *nets, int n) = 0;
    virtual void ensemble_read(int net, int kind, void *out) = 0;
    virtual void autotune(int n, int guf, bool spr, bool ang, std::vector<fn_tune_trial> &trials) = 0;
    virtual void adjoint_forward(const fn_adjoint_spec &spec, const int *group, const int *steps, int n_samples,
                                 double *samples) = 0;
    virtual void adjoint_backward(const double *dloss, double *g_spring, double *g_angle, double *g_pair,
                                  double *x0b, double *w0b) = 0;
    virtual void checkpoint_save(const char *path) = 0;
    virtual void checkpoint_load(const char *path) = 0;
    virtual void group_force(const int *group, double *out) = 0;
    virtual void group_force_noalloc(const int *group, double *out) = 0;
};

namespace {

template <typename T, typename P>
class Engine final : public fn_engine {
  public:
    explicit Engine(int dev)
    {
        device = dev;
        precision = std::is_same<T, double>::value ? FN_PREC_F64 : FN_PREC_F32;
        HIPCK(hipSetDevice(dev));
        HIPCK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        step_.alloc(1);
        stats_.alloc(fnk::kStatCount);
        HIPCK(hipHostMalloc((void **)&h_stats_, sizeof(int) * fnk::kStatCount, hipHostMallocDefault));
        fnk::fill_int(step_.p, 0, 1, stream_);
        fnk::fill_int(stats_.p, 0, fnk::kStatCount, stream_);
        for (int k = 0; k < K_COUNT; ++k) kernels_[kKernelNames[k]] = KernelRec();
        for (auto &r : role_) r = -1;
        const char *g = std::getenv("FIBRENET_GRAPHS");
        if (g && std::string(g) == "0") graphs_on_ = false;
    }
    ~Engine() override
    {
        clear_graphs();
        if (stream_) (void)hipStreamSynchronize(stream_);
        if (stream_) (void)hipStreamDestroy(stream_);
        if (h_stats_) (void)hipHostFree(h_stats_);
    }

    // ---- buffers ------------------------------------------------------------------------
    int create_buffer(int kind, int esize, int n, const void *host) override
    {
        adj_.valid = false;
        const int rs = record_size(kind);
        if (rs < 0) throw Error(FN_ERR_ARG, fmt("unknown buffer kind %d", kind));
        if (esize != rs) throw Error(FN_ERR_MISMATCH, fmt("record size %d, kind %d expects %d", esize, kind, rs));
        if (n <= 0) throw Error(FN_ERR_EMPTY, "buffer is empty");
        if (kind == FN_BUF_BEAD && (long long)n >= (1LL << fnk::kPidShift))
            throw Error(FN_ERR_UNSUPPORTED, fmt("%d beads: the engine holds fewer than 2^%d", n, fnk::kPidShift));
        const bool role_kind = kind == FN_BUF_BEAD || kind == FN_BUF_BEAD_INFO || kind == FN_BUF_SPRING ||
                               kind == FN_BUF_ANGLE || kind == FN_BUF_PAIR || kind == FN_BUF_GRID_CELL_NH;
        if (role_kind && role_[kind] >= 0)
            throw Error(FN_ERR_UNSUPPORTED, fmt("a second buffer of record kind %d", kind));
        Buf b{kind, esize, n, pad64(n)};
        const int id = (int)bufs_.size();
        bufs_.push_back(b);
        if (role_kind) role_[kind] = id;
        try {
            upload(kind, n, host);
        } catch (...) {
            bufs_.pop_back();
            if (role_kind) role_[kind] = -1;
            throw;
        }
        return id;
    }

    void write_buffer(int id, const void *host, int n_padded, int esize) override
    {
        adj_.valid = false;
        const Buf *b = buffer(id);
        if (!b) throw Error(FN_ERR_ARG, fmt("no buffer %d", id));
        if (!host || n_padded <= 0) throw Error(FN_ERR_EMPTY, "source buffer is empty");
        if (b->length != n_padded || b->esize != esize)
            throw Error(FN_ERR_MISMATCH, fmt("buffer %d holds %d x %d B, got %d x %d B", id, b->length, b->esize,
                                             n_padded, esize));
        upload(b->kind, b->elements, host);
    }

    void read_buffer(int id, void *host) override
    {
        const Buf *b = buffer(id);
        if (!b) throw Error(FN_ERR_ARG, fmt("no buffer %d", id));
        char *out = (char *)host;
        memset(out, 0, (size_t)b->length * b->esize);
        switch (b->kind) {
        case FN_BUF_BEAD: download_beads((fn_bead *)out); break;
        case FN_BUF_BEAD_INFO: memcpy(out, h_info_.data(), h_info_.size() * sizeof(fn_bead_info)); break;
        case FN_BUF_SPRING: {
            std::vector<int> st;
            sp_status_.get(st, S_, stream_);
            HIPCK(hipStreamSynchronize(stream_));
            fn_spring *o = (fn_spring *)out;
            for (int s = 0; s < S_; ++s) { o[s] = h_springs_[s]; o[s].status = st[s]; }
            break;
        }
        case FN_BUF_ANGLE: {
            std::vector<int> st;
            an_status_.get(st, A_, stream_);
            HIPCK(hipStreamSynchronize(stream_));
            fn_angle *o = (fn_angle *)out;
            for (int q = 0; q < A_; ++q) { o[q] = h_angles_[q]; o[q].status = st[q]; }
            break;
        }
        case FN_BUF_PAIR: memcpy(out, h_pairs_.data(), h_pairs_.size() * sizeof(fn_pair)); break;
        case FN_BUF_GRID_CELL_NH: memcpy(out, h_cellnh_.data(), h_cellnh_.size() * sizeof(fn_grid_cell_nh)); break;
        default:
            throw Error(FN_ERR_UNSUPPORTED, "the reference's grid/list scratch buffers are not materialised; "
                                            "use fn_export_cells / fn_export_pairs");
        }
    }

    const Buf *buffer(int id) const override
    {
        return (id >= 0 && id < (int)bufs_.size()) ? &bufs_[id] : nullptr;
    }

    // ---- kernels --------------------------------------------------------------------------
    void reg(const char *name) override
    {
        auto it = kernels_.find(name ? name : "");
        if (it == kernels_.end()) throw Error(FN_ERR_ARG, fmt("unable to create kernel: %s", name ? name : "(null)"));
        it->second.registered = true;
    }

    KernelRec &kernel(const char *name)
    {
        auto it = kernels_.find(name ? name : "");
        if (it == kernels_.end() || !it->second.registered)
            throw Error(FN_ERR_ARG, fmt("kernel not registered: %s", name ? name : "(null)"));
        return it->second;
    }

    static int kernel_id(const char *name)
    {
        for (int k = 0; k < K_COUNT; ++k)
            if (!strcmp(name, kKernelNames[k])) return k;
        return -1;
    }

    void set_arg(const char *name, int idx, int kind, const void *value, int size) override
    {
        KernelRec &kr = kernel(name);
        const int kid = kernel_id(name);
        const auto &c = kContract[kid];
        if (idx < 0 || idx >= (int)c.size()) throw Error(FN_ERR_ARG, fmt("%s has no argument %d", name, idx));
        if (c[idx].kind != kind) throw Error(FN_ERR_ARG, fmt("%s argument %d: wrong argument kind", name, idx));
        if (!value) throw Error(FN_ERR_ARG, "null argument value");
        Arg a;
        a.kind = kind;
        if (kind == FN_ARG_SETTINGS) {
            if (size != (int)sizeof(fn_settings)) throw Error(FN_ERR_MISMATCH, "settings size");
            memcpy(&a.s, value, sizeof(fn_settings));
        } else {
            if (size != (int)sizeof(int)) throw Error(FN_ERR_MISMATCH, "argument size");
            a.ival = *(const int *)value;
            if (kind == FN_ARG_BUFFER) {
                const Buf *b = buffer(a.ival);
                if (!b || b->kind != c[idx].buf)
                    throw Error(FN_ERR_ARG, fmt("%s argument %d: buffer %d is not of record kind %d", name, idx,
                                                a.ival, c[idx].buf));
            }
        }
        kr.args[idx] = a;
        clear_graphs();
    }

    void set_range(const char *name, int range) override { kernel(name).range = range; }

    // Effective count of a launch: the bound count clipped to the work range.
    int active(int kid, const KernelRec &kr)
    {
        const auto &c = kContract[kid];
        for (int i = 0; i < (int)c.size(); ++i)
            if (!kr.args.count(i)) throw Error(FN_ERR_ARG, fmt("%s: argument %d not set", kKernelNames[kid], i));
        if (kr.range <= 0) throw Error(FN_ERR_ARG, fmt("%s: incorrect kernel work range", kKernelNames[kid]));
        return std::min(kr.args.at(kCountArg[kid]).ival, kr.range);
    }

    const fn_settings &settings_of(int kid)
    {
        return kernels_.at(kKernelNames[kid]).args.at(kSettingsArg[kid]).s;
    }

    void launch(const char *name) override
    {
        adj_.valid = false;
        KernelRec &kr = kernel(name);
        const int kid = kernel_id(name);
        const int n = active(kid, kr);
        prepare();
        switch (kid) {
        case K_ADVANCE_POSITION: {
            auto a = args_granular();
            a.drift = make_box(kr.args.at(2).s);
            check_alias(a.drift.bc);
            a = args_granular();
            a.drift = make_box(kr.args.at(2).s);
            fnk::Launch<T, P>::advance_position(a, std::min(n, N_), stream_);
            break;
        }
        case K_RESOLVE_SPRINGS: {
            auto a = args_granular();
            a.n_springs = std::min(n, S_);
            if (S_ > 0 && a.n_springs > 0) fnk::Launch<T, P>::resolve_springs(a, S_, stream_);
            break;
        }
        case K_RESOLVE_ANGLES: {
            auto a = args_granular();
            a.n_angles = std::min(n, A_);
            if (A_ > 0 && a.n_angles > 0) fnk::Launch<T, P>::resolve_angles(a, A_, stream_);
            break;
        }
        case K_RESET_GRID: grid_valid_ = false; break;
        case K_RESET_LIST:
            ensure_grid(nullptr);
            fnk::fill_int(nb_cnt_.p, 0, N_, stream_);
            break;
        case K_FILL_GRID:
            if (n != N_) throw Error(FN_ERR_UNSUPPORTED, "FillGrid over a partial bead range");
            hash_and_sort(make_box(kr.args.at(3).s));
            break;
        case K_BUILD_LIST: {
            const fn_settings &s = kr.args.at(9).s;
            if (n != s.cells_x * s.cells_y * s.cells_z)
                throw Error(FN_ERR_UNSUPPORTED, "BuildContactList over a partial grid");
            if (!grid_valid_) throw Error(FN_ERR_ARG, "BuildContactList before FillGrid");
            build_list(make_box(s));
            break;
        }
        case K_RESOLVE_CONTACTS: {
            if (n != N_) throw Error(FN_ERR_UNSUPPORTED, "ResolveContacts over a partial bead range");
            ensure_grid(nullptr);
            auto a = args_granular();
            a.force = make_box(kr.args.at(6).s);
            fnk::Launch<T, P>::resolve_contacts(a, stream_);
            break;
        }
        case K_ADVANCE_VELOCITY: {
            auto a = args_granular();
            fnk::Launch<T, P>::advance_velocity(a, std::min(n, N_), stream_);
            ++steps_;
            break;
        }
        case K_RESET_ENERGY: {
            auto a = args_granular();
            fnk::Launch<T, P>::reset_energy(a, std::min(n, N_), stream_);
            break;
        }
        }
        HIPCK(hipGetLastError());
    }

    void finish() override { HIPCK(hipStreamSynchronize(stream_)); }

    // ---- fast path --------------------------------------------------------------------------
    void run_steps(int n, int guf, bool spr, bool ang) override
    {
        adj_.valid = false;
        prep_fast(n, guf, spr, ang);
        presize_rows();
        if (rows_exact_ > 0 && (!keep_cand_ || rows_cand_ > 0)) {
            run_chunks(n, guf, spr, ang);
            return;
        }
        // No search-row lengths seen yet for this state (or for kept
        // candidates): keep the state, so that a run whose search rows or
        // contact list overflow is repeated with the grown capacities.
        const std::vector<Part> parts = state_parts(pc_, fc_, false);
        run_snap_.ensure(parts_bytes(parts));
        save_parts(parts, run_snap_.p);
        const int pc = pc_, fc = fc_;
        const long long steps = steps_, rebuilds = rebuilds_;
        for (int attempt = 0;; ++attempt) {
            try {
                run_chunks(n, guf, spr, ang);
                return;
            } catch (const Error &e) {
                if (e.code != FN_ERR_CAPACITY || attempt == 3) throw;
            }
            pc_ = pc;
            fc_ = fc;
            steps_ = steps;
            rebuilds_ = rebuilds;
            load_parts(state_parts(pc_, fc_, false), run_snap_.p);
            fnk::fill_int(need_.p, 1, 1, stream_);      // kept candidates are built afresh
        }
    }

    // Search rows long enough for the longest rows seen so far (exact lists,
    // and kept candidates when they are in use), with room to grow.
    void presize_rows()
    {
        const int need = std::max(rows_exact_, keep_cand_ ? rows_cand_ : 0);
        if (need * 3LL / 2 <= K_) return;
        long long k = K_;
        while (k < 2LL * need) k *= 2;
        set_rows(k);
        ensure_scratch();
    }

    // Checks and device set-up of a fast-path run; spr / ang become the bond
    // passes that run.
    void prep_fast(int n, int guf, bool &spr, bool &ang)
    {
        if (n <= 0) throw Error(FN_ERR_ARG, "loop size should be a positive number");
        if (guf <= 0) throw Error(FN_ERR_ARG, "grid_update_freq should be positive");
        if (!ens_boxes_.empty()) {
            prepare_ensemble_run(spr, ang);
            spr = spr && S_ > 0;
            ang = ang && A_ > 0;
            return;
        }
        for (int kid : {K_ADVANCE_POSITION, K_FILL_GRID, K_BUILD_LIST, K_RESOLVE_CONTACTS, K_ADVANCE_VELOCITY}) {
            const KernelRec &kr = kernel(kKernelNames[kid]);
            const int m = active(kid, kr);
            const int want = kid == K_BUILD_LIST ? C_of(kr.args.at(9).s) : N_;
            if (m != want) throw Error(FN_ERR_UNSUPPORTED, "fast path needs full-range kernels");
        }
        if (spr) {
            const KernelRec &kr = kernel("ResolveSprings");
            n_spr_ = std::min(active(K_RESOLVE_SPRINGS, kr), S_);
        }
        if (ang) {
            const KernelRec &kr = kernel("ResolveAngles");
            n_ang_ = std::min(active(K_RESOLVE_ANGLES, kr), A_);
        }
        spr = spr && S_ > 0 && n_spr_ > 0;
        ang = ang && A_ > 0 && n_ang_ > 0;
        prepare();
        check_alias(settings_of(K_ADVANCE_POSITION).bc_type);
        ensure_grid(&settings_of(K_FILL_GRID));
        ensure_grid(&settings_of(K_BUILD_LIST));
        ensure_scratch();
        geometry(settings_of(K_FILL_GRID), settings_of(K_BUILD_LIST));
        if (cand_dirty_) {
            fnk::fill_int(need_.p, 1, 1, stream_);
            cand_dirty_ = false;
        }
    }

    // The chunks of a fast-path run, then the checks after it.
    void run_chunks(int n, int guf, bool spr, bool ang)
    {
        // capacity of the contact list: grow before the run if the last rebuild came close
        if ((long long)last_total_ * 3 / 2 + 1024 > (long long)cap_) grow_list((long long)last_total_ * 2 + 1024);
        fnk::fill_int(stats_.p, 0, fnk::kStatCount, stream_);
        const int G = chunk_len(guf);
        int done = 0;
        while (done < n) {
            const int len = std::min(G, n - done);
            run_chunk(len, guf, spr, ang, done == 0, done + len == n);
            done += len;
        }
        steps_ += n;
        const int rb = (n + guf - 1) / guf;
        rebuilds_ += rb;
        // the run's statistics come back with the stream's last work
        HIPCK(hipMemcpyAsync(h_stats_, stats_.p, sizeof(int) * fnk::kStatCount, hipMemcpyDeviceToHost, stream_));
        HIPCK(hipStreamSynchronize(stream_));
        collect_timing();
        check_capacity(h_stats_);
        tune_verlet(rb);
    }

    // ---- exports --------------------------------------------------------------------------------
    void export_cells(int *out) override
    {
        if (!grid_valid_) throw Error(FN_ERR_ARG, "no grid has been built");
        HIPCK(hipMemcpyAsync(out, key_.p, sizeof(int) * N_, hipMemcpyDeviceToHost, stream_));
        HIPCK(hipStreamSynchronize(stream_));
    }

    long long export_pairs(int *out, long long cap) override
    {
        std::vector<int> st, cnt, lst;
        nb_start_.get(st, N_ + 1, stream_);
        nb_cnt_.get(cnt, N_, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        const int used = std::min(st[N_], cap_);
        lst.resize(used);
        if (used) HIPCK(hipMemcpyAsync(lst.data(), nb_list_.p, sizeof(int) * used, hipMemcpyDeviceToHost, stream_));
        HIPCK(hipStreamSynchronize(stream_));
        std::vector<long long> keys;
        for (int i = 0; i < N_; ++i)
            for (int e = st[i]; e < st[i] + cnt[i] && e < used; ++e) {
                const int j = lst[e] & fnk::kIdxMask;
                if (i < j) keys.push_back((long long)i * N_ + j);
            }
        std::sort(keys.begin(), keys.end());
        if (out) {
            const long long m = std::min<long long>(cap, (long long)keys.size());
            for (long long k = 0; k < m; ++k) {
                out[2 * k] = (int)(keys[k] / N_);
                out[2 * k + 1] = (int)(keys[k] % N_);
            }
        }
        return (long long)keys.size();
    }

    void export_breaks(int *out) override
    {
        if (S_) HIPCK(hipMemcpyAsync(out, sp_break_.p, sizeof(int) * S_, hipMemcpyDeviceToHost, stream_));
        HIPCK(hipStreamSynchronize(stream_));
    }

    void stats(fn_stats *o) override
    {
        int st[fnk::kStatCount];
        HIPCK(hipMemcpyAsync(st, stats_.p, sizeof st, hipMemcpyDeviceToHost, stream_));
        int cand = 0;
        if (nb_start_.p) HIPCK(hipMemcpyAsync(&cand, nb_start_.p + N_, sizeof(int), hipMemcpyDeviceToHost, stream_));
        std::vector<int> ss, keys, cnt;
        if (key_.p) key_.get(keys, N_, stream_);
        if (nb_cnt_.p) nb_cnt_.get(cnt, N_, stream_);
        sp_status_.get(ss, S_, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        memset(o, 0, sizeof *o);
        o->steps = steps_;
        o->rebuilds = rebuilds_;
        long long listed = 0;
        for (int c : cnt) listed += c;
        o->list_entries = listed;
        o->list_capacity = cap_;
        o->candidate_entries = cand;
        o->candidate_builds = st[fnk::kStatCandBuilds];
        o->hashed_beads = (int)std::count_if(keys.begin(), keys.end(), [](int k) { return k >= 0; });
        o->max_cell_occupancy = st[fnk::kStatMaxCell];
        o->cells_over_capacity = st[fnk::kStatCellsOver];
        o->max_list_length = st[fnk::kStatMaxList];
        o->broken_springs = (int)std::count(ss.begin(), ss.end(), 0);
        o->img_aliases_pos = alias_ ? 1 : 0;
    }

    void set_timing(bool on) override
    {
        if (on != timing_) clear_graphs();
        timing_ = on;
        timing_acc_ = fn_timing{};
    }

    void timing(fn_timing *out) override { *out = timing_acc_; }

    void copy_probe(long long bytes, double *ms) override
    {
        const size_t n16 = ((size_t)bytes + 15) / 16;
        DArr<int> src, dst;
        src.alloc(n16 * 4);
        dst.alloc(n16 * 4);
        fnk::fill_int(src.p, 1, (int)std::min<size_t>(n16 * 4, (size_t)INT32_MAX), stream_);
        hipEvent_t e0 = new_event(), e1 = new_event();
        HIPCK(hipEventRecord(e0, stream_));
        fnk::copy_probe(src.p, dst.p, n16, stream_);
        HIPCK(hipEventRecord(e1, stream_));
        HIPCK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCK(hipEventElapsedTime(&t, e0, e1));
        *ms = t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIPCK(hipStreamSynchronize(stream_));
    }

    void set_option(const std::string &name, int value) override
    {
        if (name == "chunk") chunk_ = value;
        else if (name == "chain") chain_on_ = value != 0;
        else if (name == "staged") staged_ = value != 0;
        else if (name == "flat_search") flat_ = value != 0;
        else if (name == "verlet") {
            if (value < 0 || value > 2) throw Error(FN_ERR_ARG, "verlet is 0 (off), 1 (on) or 2 (auto)");
            verlet_ = value;
            verlet_trial_ = false;
            verlet_runs_ = 0;
        }
        else if (name == "skin") {
            if (value < 0 || value > 1000) throw Error(FN_ERR_ARG, "skin must lie in [0, 1000] permille of the reach");
            if (value != skin_permille_) adj_.valid = false;   // its recomputation keeps the run's candidates
            skin_permille_ = value;
        }
        else if (name == "graphs") graphs_on_ = value != 0;
        else if (name == "block") {
            if (value != 128 && value != 256 && value != 512)
                throw Error(FN_ERR_ARG, "block is 128, 256 or 512 threads");
            block_ = value;
        }
        else throw Error(FN_ERR_ARG, "unknown option " + name);
        clear_graphs();
    }

    int get_option(const std::string &name) override
    {
        if (name == "chunk") return chunk_;
        if (name == "chain") return chain_on_ ? 1 : 0;
        if (name == "staged") return staged_ ? 1 : 0;
        if (name == "flat_search") return flat_ ? 1 : 0;
        if (name == "verlet") return verlet_;
        if (name == "skin") return skin_permille_;
        if (name == "chain_layout") {
            prepare();
            return chain_ok_ ? 1 : 0;
        }
        if (name == "chain_sets") {
            prepare();
            return chain_ok_ && cmp_ok_ ? 1 : 0;
        }
        if (name == "graphs") return graphs_on_ ? 1 : 0;
        if (name == "block") return block_;
        throw Error(FN_ERR_ARG, "unknown option " + name);
    }

  private:
    static int C_of(const fn_settings &s) { return s.cells_x * s.cells_y * s.cells_z; }

    // Events of one chunk: four per step (before the spring terms, before the
    // angle terms, before the bead kernel, after it) and a pair per rebuild.
    struct Events {
        std::vector<hipEvent_t> step, rebuild;
        void release()
        {
            for (auto e : step) (void)hipEventDestroy(e);
            for (auto e : rebuild) (void)hipEventDestroy(e);
            step.clear();
            rebuild.clear();
        }
    };

    static hipEvent_t new_event()
    {
        hipEvent_t e;
        HIPCK(hipEventCreate(&e));
        return e;
    }

    // Records a timing event on the stream; under graph capture as an
    // external event node, which records each time the graph runs.
    void mark(std::vector<hipEvent_t> *v)
    {
        if (!v) return;
        v->push_back(new_event());
        if (capturing_) HIPCK(hipEventRecordWithFlags(v->back(), stream_, hipEventRecordExternal));
        else HIPCK(hipEventRecord(v->back(), stream_));
    }

    void accumulate(const Events &ev)
    {
        for (size_t k = 0; k + 3 < ev.step.size(); k += 4) {
            float ms[3] = {0.f, 0.f, 0.f};
            for (int j = 0; j < 3; ++j) HIPCK(hipEventElapsedTime(&ms[j], ev.step[k + j], ev.step[k + j + 1]));
            timing_acc_.spring_ms += ms[0];
            timing_acc_.angle_ms += ms[1];
            timing_acc_.bead_ms += ms[2];
            timing_acc_.step_ms += (double)ms[0] + ms[1] + ms[2];
            timing_acc_.steps++;
        }
        for (size_t k = 0; k + 1 < ev.rebuild.size(); k += 2) {
            float ms = 0.f;
            HIPCK(hipEventElapsedTime(&ms, ev.rebuild[k], ev.rebuild[k + 1]));
            timing_acc_.rebuild_ms += ms;
            timing_acc_.rebuilds++;
        }
    }

    // ---- upload / download -------------------------------------------------------------------
    template <typename X>
    static void put3(DArr<X> *d, const std::vector<X> &x, const std::vector<X> &y, const std::vector<X> &z,
                     hipStream_t s)
    {
        d[0].put(x, s);
        d[1].put(y, s);
        d[2].put(z, s);
    }

    void alloc_beads(int n)
    {
        clear_graphs();
        N_ = n;
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 3; ++c) {
                pos_[b][c].alloc(n);
                frc_[b][c].alloc(n);
                img_[b][c].release();
            }
        for (int c = 0; c < 3; ++c) {
            vel_[c].alloc(n);
            vold_[c].alloc(n);
        }
        lam_.alloc(n); hlf_.alloc(n); hlf2_.alloc(n); rad_.alloc(n);
        type_.alloc(n); tag_.alloc(n); loc_.alloc(n);
        rad_d_.alloc(n); skin_d_.alloc(n);
        alias_ = false;
        csr_dirty_ = true;
        grid_valid_ = false;
    }

    void upload(int kind, int n, const void *host)
    {
        HIPCK(hipSetDevice(device));
        switch (kind) {
        case FN_BUF_BEAD: {
            std::vector<fn_bead> z;
            const fn_bead *b = (const fn_bead *)host;
            if (!b) { z.assign(n, fn_bead{}); b = z.data(); }
            if (n != N_) alloc_beads(n);
            upload_beads(b);
            break;
        }
        case FN_BUF_BEAD_INFO: {
            if (role_[FN_BUF_BEAD] < 0 || n != N_)
                throw Error(FN_ERR_ARG, "BeadInfo buffer must follow the Bead buffer and match its length");
            h_info_.assign(n, fn_bead_info{});
            if (host) memcpy(h_info_.data(), host, sizeof(fn_bead_info) * n);
            upload_info();
            break;
        }
        case FN_BUF_SPRING:
            h_springs_.assign(n, fn_spring{});
            if (host) memcpy(h_springs_.data(), host, sizeof(fn_spring) * n);
            S_ = n;
            upload_springs();
            break;
        case FN_BUF_ANGLE:
            h_angles_.assign(n, fn_angle{});
            if (host) memcpy(h_angles_.data(), host, sizeof(fn_angle) * n);
            A_ = n;
            upload_angles();
            break;
        case FN_BUF_PAIR: {
            h_pairs_.assign(n, fn_pair{});
            if (host) memcpy(h_pairs_.data(), host, sizeof(fn_pair) * n);
            std::vector<fnk::PairC> pc(n);
            for (int k = 0; k < n; ++k) {
                const fn_pair &p = h_pairs_[k];
                pc[k] = fnk::PairC{p.type, p.c1, p.c2, p.c3, p.c4, p.c5, p.c6};
            }
            pairs_.put(pc, stream_);
            cand_dirty_ = true;
            rows_exact_ = rows_cand_ = 0;
            clear_graphs();
            break;
        }
        case FN_BUF_GRID_CELL_NH:
            h_cellnh_.assign(n, fn_grid_cell_nh{});
            if (host) memcpy(h_cellnh_.data(), host, sizeof(fn_grid_cell_nh) * n);
            cellnh_checked_ = false;
            break;
        default: break;   // reference scratch (grid heads/tails, list heads/tails): not materialised
        }
        HIPCK(hipStreamSynchronize(stream_));
    }

    void upload_beads(const fn_bead *b)
    {
        const int n = N_;
        std::vector<P> x(n), y(n), z(n), ix(n), iy(n), iz(n);
        std::vector<T> a(n), bb(n), c(n);
        bool eq = true;
        for (int i = 0; i < n; ++i) {
            x[i] = (P)b[i].position.x; y[i] = (P)b[i].position.y; z[i] = (P)b[i].position.z;
            ix[i] = (P)b[i].image_position.x; iy[i] = (P)b[i].image_position.y; iz[i] = (P)b[i].image_position.z;
            eq = eq && x[i] == ix[i] && y[i] == iy[i] && z[i] == iz[i];
        }
        pc_ = 0;
        fc_ = 0;
        put3(pos_[0], x, y, z, stream_);
        pos_eq_img_ = eq;
        if (alias_ && !eq) unalias();
        if (!alias_) {
            for (int k = 0; k < 2; ++k)
                for (int q = 0; q < 3; ++q) img_[k][q].ensure(n);
            put3(img_[0], ix, iy, iz, stream_);
        }
        auto field = [&](DArr<T> *d, const fn_vector3 fn_bead::*m) {
            for (int i = 0; i < n; ++i) { a[i] = (T)(b[i].*m).x; bb[i] = (T)(b[i].*m).y; c[i] = (T)(b[i].*m).z; }
            put3(d, a, bb, c, stream_);
            HIPCK(hipStreamSynchronize(stream_));
        };
        field(vel_, &fn_bead::velocity);
        field(vold_, &fn_bead::velocity_old);
        field(frc_[0], &fn_bead::force);
        field(frc_[1], &fn_bead::force_old);
        HIPCK(hipStreamSynchronize(stream_));
        cand_dirty_ = true;
        rows_exact_ = rows_cand_ = 0;
    }

    void download_beads(fn_bead *o)
    {
        const int n = N_;
        std::vector<P> x, y, z, ix, iy, iz;
        auto get3p = [&](DArr<P> *d, std::vector<P> &a, std::vector<P> &b, std::vector<P> &c) {
            d[0].get(a, n, stream_); d[1].get(b, n, stream_); d[2].get(c, n, stream_);
        };
        get3p(pos_[pc_], x, y, z);
        get3p(alias_ ? pos_[pc_] : img_[pc_], ix, iy, iz);
        std::vector<T> t[4][3];
        DArr<T> *src[4] = {vel_, frc_[fc_], vold_, frc_[1 - fc_]};
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c) src[k][c].get(t[k][c], n, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        for (int i = 0; i < n; ++i) {
            fn_bead &b = o[i];
            b.position = fn_vector3{(double)x[i], (double)y[i], (double)z[i], 0.0};
            b.image_position = fn_vector3{(double)ix[i], (double)iy[i], (double)iz[i], 0.0};
            fn_vector3 *dst[4] = {&b.velocity, &b.force, &b.velocity_old, &b.force_old};
            for (int k = 0; k < 4; ++k)
                *dst[k] = fn_vector3{(double)t[k][0][i], (double)t[k][1][i], (double)t[k][2][i], 0.0};
        }
    }

    void upload_info()
    {
        const int n = N_;
        std::vector<T> l(n), h(n), h2(n), r(n);
        std::vector<int> ty(n), tg(n), lc(n);
        std::vector<double> rd(n), sk(n);
        for (int i = 0; i < n; ++i) {
            const fn_bead_info &q = h_info_[i];
            l[i] = (T)q.lambda_dt_mass_inv;
            h[i] = (T)q.half_dt_mass_inv;
            h2[i] = (T)q.half_dt2_mass_inv;
            r[i] = (T)q.radius;
            ty[i] = q.type;
            tg[i] = q.tag;
            lc[i] = q.loc_id;
            rd[i] = q.radius;
            sk[i] = q.skin_radius;
        }
        lam_.put(l, stream_); hlf_.put(h, stream_); hlf2_.put(h2, stream_); rad_.put(r, stream_);
        build_mass_sets(l, h, h2, r);
        type_.put(ty, stream_); tag_.put(tg, stream_); loc_.put(lc, stream_);
        rad_d_.put(rd, stream_); skin_d_.put(sk, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        max_radius_ = 0.0;
        max_skin_ = 0.0;
        types_checked_ = -1;
        for (const fn_bead_info &q : h_info_) {
            max_radius_ = std::max(max_radius_, q.radius);
            max_skin_ = std::max(max_skin_, q.skin_radius);
        }
        cand_dirty_ = true;
        rows_exact_ = rows_cand_ = 0;
        clear_graphs();
    }

    void upload_springs()
    {
        const int n = S_;
        std::vector<int> p1(n), p2(n), ty(n), st(n), br(n, -1);
        std::vector<T> k(n);
        std::vector<double> l0(n), c1(n), c2(n);
        for (int s = 0; s < n; ++s) {
            const fn_spring &q = h_springs_[s];
            p1[s] = q.p1; p2[s] = q.p2; ty[s] = q.type; st[s] = q.status;
            l0[s] = q.rest_length; k[s] = (T)q.stiffness; c1[s] = q.c1; c2[s] = q.c2;
        }
        sp_p1_.put(p1, stream_); sp_p2_.put(p2, stream_); sp_type_.put(ty, stream_); sp_status_.put(st, stream_);
        sp_break_.put(br, stream_);
        sp_l0_.put(l0, stream_); sp_k_.put(k, stream_); sp_c1_.put(c1, stream_); sp_c2_.put(c2, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        csr_dirty_ = true;
        clear_graphs();
    }

    void upload_angles()
    {
        const int n = A_;
        std::vector<int> b1(n), b2(n), b3(n), s1(n), s2(n), ty(n), st(n);
        std::vector<T> th(n), k(n);
        for (int q = 0; q < n; ++q) {
            const fn_angle &a = h_angles_[q];
            b1[q] = a.b1; b2[q] = a.b2; b3[q] = a.b3; s1[q] = a.s1; s2[q] = a.s2; ty[q] = a.type; st[q] = a.status;
            th[q] = (T)a.angle; k[q] = (T)a.c1;
        }
        an_b1_.put(b1, stream_); an_b2_.put(b2, stream_); an_b3_.put(b3, stream_);
        an_s1_.put(s1, stream_); an_s2_.put(s2, stream_); an_type_.put(ty, stream_); an_status_.put(st, stream_);
        an_th0_.put(th, stream_); an_k_.put(k, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        csr_dirty_ = true;
        clear_graphs();
    }

    // Incidence slots bead -> springs and bead -> angles, ordered by bond
    // index (then p1 before p2; b1, b3, b2) as the reference's sequential
    // scatter adds them: the bond kernels write each term to its slot, the
    // bead kernels sum a bead's slots in order.
    void build_csr()
    {
        const int n = N_;
        std::vector<int> cnt(n + 1, 0);
        for (const fn_spring &s : h_springs_) {
            if (s.p1 < 0 || s.p1 >= n || s.p2 < 0 || s.p2 >= n) throw Error(FN_ERR_ARG, "spring bead index out of range");
            cnt[s.p1]++; cnt[s.p2]++;
        }
        std::vector<int> st(n + 1, 0), slot(2 * (size_t)S_);
        for (int i = 0; i < n; ++i) st[i + 1] = st[i] + cnt[i];
        std::vector<int> cur(st.begin(), st.end() - 1);
        for (int s = 0; s < S_; ++s) {
            slot[2 * (size_t)s] = cur[h_springs_[s].p1]++;
            slot[2 * (size_t)s + 1] = cur[h_springs_[s].p2]++;
        }
        sp_start_.put(st, stream_);
        sp_slot_.put(slot, stream_);
        sp_term_.ensure(2 * (size_t)S_);
        std::fill(cnt.begin(), cnt.end(), 0);
        for (const fn_angle &a : h_angles_) {
            for (int b : {a.b1, a.b2, a.b3})
                if (b < 0 || b >= n) throw Error(FN_ERR_ARG, "angle bead index out of range");
            if (a.type == 1 && (a.s1 < 0 || a.s1 >= S_ || a.s2 < 0 || a.s2 >= S_))
                throw Error(FN_ERR_ARG, "HR_FL angle spring index out of range");
            cnt[a.b1]++; cnt[a.b2]++; cnt[a.b3]++;
        }
        std::vector<int> as(n + 1, 0), aslot(3 * (size_t)A_);
        for (int i = 0; i < n; ++i) as[i + 1] = as[i] + cnt[i];
        std::vector<int> acur(as.begin(), as.end() - 1);
        for (int q = 0; q < A_; ++q) {
            aslot[3 * (size_t)q] = acur[h_angles_[q].b1]++;
            aslot[3 * (size_t)q + 1] = acur[h_angles_[q].b3]++;
            aslot[3 * (size_t)q + 2] = acur[h_angles_[q].b2]++;
        }
        an_start_.put(as, stream_);
        an_slot_.put(aslot, stream_);
        an_term_.ensure(3 * (size_t)A_);
        build_chain();
        HIPCK(hipStreamSynchronize(stream_));
        csr_dirty_ = false;
        clear_graphs();
    }

    // Chain layout: every spring joins beads (i, i+1) and every angle is
    // (i-1, i, i+1), at most one of each per bead, numbered along the chain
    // (spring (i-1, i) before (i, i+1); angle centred on i-1 before the one on
    // i before the one on i+1, the order of each bead's incidence list), and
    // an HR_FL angle fails with the two chain springs of its centre bead.
    void build_chain()
    {
        const int n = N_;
        std::vector<int> cs(n, -1), ca(n, -1);
        bool ok = true;
        for (int s = 0; s < S_ && ok; ++s) {
            const fn_spring &q = h_springs_[s];
            if (q.p2 != q.p1 + 1 || cs[q.p1] >= 0) ok = false;
            else cs[q.p1] = s;
        }
        for (int q = 0; q < A_ && ok; ++q) {
            const fn_angle &g = h_angles_[q];
            if (g.b2 != g.b1 + 1 || g.b3 != g.b2 + 1 || ca[g.b2] >= 0) ok = false;
            else ca[g.b2] = q;
        }
        for (int i = 0; i < n && ok; ++i) {
            if (i > 0 && cs[i - 1] >= 0 && cs[i] >= 0 && cs[i - 1] > cs[i]) ok = false;
            int last = -1;
            for (int k = i - 1; k <= i + 1; ++k)
                if (k >= 0 && k < n && ca[k] >= 0) {
                    if (ca[k] < last) ok = false;
                    last = ca[k];
                }
            if (ca[i] >= 0 && h_angles_[ca[i]].type == 1) {
                const fn_angle &g = h_angles_[ca[i]];
                const int l = i > 0 ? cs[i - 1] : -1, r = cs[i];
                if (l < 0 || r < 0 || !((g.s1 == l && g.s2 == r) || (g.s1 == r && g.s2 == l))) ok = false;
            }
        }
        chain_ok_ = ok;
        if (ok) {
            ch_spring_.put(cs, stream_);
            ch_angle_.put(ca, stream_);
        }
        build_chain_sets(cs, ca);
    }

    // The beads' mass coefficients and radii as a one-byte code per bead into
    // their distinct sets, when there are at most kMassSets of them (the
    // chain kernel reads the code in place of four values per bead).
    void build_mass_sets(const std::vector<T> &l, const std::vector<T> &h, const std::vector<T> &h2,
                         const std::vector<T> &r)
    {
        ms_ok_ = false;
        std::map<std::tuple<T, T, T, T>, int> key;
        std::vector<fnk::MassSet<T>> sets;
        std::vector<unsigned char> code(N_);
        for (int i = 0; i < N_; ++i) {
            const auto k = std::make_tuple(l[i], h[i], h2[i], r[i]);
            auto it = key.find(k);
            if (it == key.end()) {
                if ((int)sets.size() == fnk::kMassSets) return;
                it = key.emplace(k, (int)sets.size()).first;
                sets.push_back(fnk::MassSet<T>{l[i], h[i], h2[i], r[i]});
            }
            code[i] = (unsigned char)it->second;
        }
        ms_code_.put(code, stream_);
        ms_sets_.put(sets, stream_);
        n_ms_ = (int)sets.size();
        HIPCK(hipStreamSynchronize(stream_));
        ms_ok_ = true;
    }

    // The chain layout's bond parameters by bead: one word per bead with the
    // codes of its spring's distinct (type, l0, E, failure stress, k) set and
    // its angle's (type, theta0, k) set, when each kind has at most kNoSet
    // sets; the status bits are filled by chain_status before every chunk.
    void build_chain_sets(const std::vector<int> &cs, const std::vector<int> &ca)
    {
        cmp_ok_ = false;
        if (!chain_ok_) return;
        const int n = N_;
        std::map<std::tuple<int, double, double, double, double>, int> sk;
        std::map<std::tuple<int, double, double>, int> ak;
        std::vector<fnk::SpringSet> ss;
        std::vector<fnk::AngleSet> as;
        std::vector<unsigned> code(n, fnk::kNoSet | (fnk::kNoSet << 16));
        for (int i = 0; i < n; ++i) {
            unsigned sc = fnk::kNoSet, ac = fnk::kNoSet;
            if (cs[i] >= 0) {
                const fn_spring &q = h_springs_[cs[i]];
                if (q.status != 0 && q.status != 1) return;   // a status bit holds 0 or 1 only
                const auto key = std::make_tuple(q.type, q.rest_length, q.c1, q.c2, (double)(T)q.stiffness);
                auto it = sk.find(key);
                if (it == sk.end()) {
                    if ((unsigned)ss.size() >= fnk::kNoSet) return;
                    it = sk.emplace(key, (int)ss.size()).first;
                    ss.push_back(fnk::SpringSet{q.type, q.rest_length, q.c1, q.c2, (double)(T)q.stiffness});
                }
                sc = (unsigned)it->second;
            }
            if (ca[i] >= 0) {
                const fn_angle &g = h_angles_[ca[i]];
                if (g.status != 0 && g.status != 1) return;
                const auto key = std::make_tuple(g.type, (double)(T)g.angle, (double)(T)g.c1);
                auto it = ak.find(key);
                if (it == ak.end()) {
                    if ((unsigned)as.size() >= fnk::kNoSet) return;
                    it = ak.emplace(key, (int)as.size()).first;
                    as.push_back(fnk::AngleSet{g.type, (double)(T)g.angle, (double)(T)g.c1});
                }
                ac = (unsigned)it->second;
            }
            code[i] = sc | (ac << 16);
        }
        bond_code_.put(code, stream_);
        sp_sets_.put(ss.empty() ? std::vector<fnk::SpringSet>(1) : ss, stream_);
        an_sets_.put(as.empty() ? std::vector<fnk::AngleSet>(1) : as, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        cmp_ok_ = true;
    }

    void prepare()
    {
        if (N_ <= 0 || role_[FN_BUF_BEAD_INFO] < 0) throw Error(FN_ERR_ARG, "bead and bead-info buffers are required");
        if (role_[FN_BUF_PAIR] < 0) throw Error(FN_ERR_ARG, "pair table buffer is required");
        if (csr_dirty_) build_csr();
        if ((long long)N_ >= (1LL << fnk::kPidShift) ||
            (ens_boxes_.empty() && (int)h_pairs_.size() > (1 << (31 - fnk::kPidShift))))
            throw Error(FN_ERR_UNSUPPORTED, "more than 2^26 beads or 32 type pairs");
    }

    // One array holds position and image when they are equal and the box is
    // finite: both then receive the same increments forever (no periodic
    // reset), so the reference's two fields stay bit-identical.
    void check_alias(int bc)
    {
        const bool want = bc == 1 && pos_eq_img_;
        if (want == alias_) return;
        clear_graphs();
        if (want) {
            for (int k = 0; k < 2; ++k)
                for (int q = 0; q < 3; ++q) img_[k][q].release();
            alias_ = true;
        } else {
            unalias();
        }
    }

    void unalias()
    {
        clear_graphs();
        for (int k = 0; k < 2; ++k)
            for (int q = 0; q < 3; ++q) {
                img_[k][q].ensure(N_);
                if (alias_) HIPCK(hipMemcpyAsync(img_[k][q].p, pos_[k][q].p, sizeof(P) * N_, hipMemcpyDeviceToDevice, stream_));
            }
        HIPCK(hipStreamSynchronize(stream_));
        alias_ = false;
        pos_eq_img_ = false;
    }

    fnk::Args<T, P> args_base()
    {
        fnk::Args<T, P> a{};
        a.n = N_;
        a.vx = vel_[0].p; a.vy = vel_[1].p; a.vz = vel_[2].p;
        a.wx = vold_[0].p; a.wy = vold_[1].p; a.wz = vold_[2].p;
        a.lam = lam_.p; a.hlf = hlf_.p; a.hlf2 = hlf2_.p;
        a.type = type_.p; a.tag = tag_.p; a.loc = loc_.p; a.radius = rad_.p;
        a.sp_start = sp_start_.p; a.sp_p1 = sp_p1_.p; a.sp_p2 = sp_p2_.p;
        a.sp_type = sp_type_.p; a.sp_status = sp_status_.p; a.sp_break = sp_break_.p;
        a.sp_l0 = sp_l0_.p; a.sp_k = sp_k_.p; a.sp_c1 = sp_c1_.p; a.sp_c2 = sp_c2_.p;
        a.n_springs = S_;
        a.an_start = an_start_.p; a.an_b1 = an_b1_.p; a.an_b2 = an_b2_.p; a.an_b3 = an_b3_.p;
        a.an_s1 = an_s1_.p; a.an_s2 = an_s2_.p; a.an_type = an_type_.p; a.an_status = an_status_.p;
        a.an_th0 = an_th0_.p; a.an_k = an_k_.p;
        a.n_angles = A_;
        a.nb_start = nb_start_.p; a.nb_cnt = nb_cnt_.p; a.nb_list = nb_list_.p; a.pairs = pairs_.p;
        a.n_pairs = (int)h_pairs_.size();
        a.sp_term = sp_term_.p; a.an_term = an_term_.p;
        a.sp_slot = sp_slot_.p; a.an_slot = an_slot_.p;
        a.ch_spring = ch_spring_.p; a.ch_angle = ch_angle_.p;
        if (chain_ok_ && cmp_ok_) {
            a.bond_code = bond_code_.p;
            a.sp_sets = sp_sets_.p; a.an_sets = an_sets_.p;
        }
        if (ms_ok_ && ens_boxes_.empty()) {
            a.ms_code = ms_code_.p; a.ms_sets = ms_sets_.p; a.n_ms = n_ms_;
        }
        a.net = ens_net();
        a.boxes = boxes_.p;
        a.step = step_.p;
        return a;
    }

    void set_pos(fnk::Args<T, P> &a, int cur)
    {
        a.px = pos_[cur][0].p; a.py = pos_[cur][1].p; a.pz = pos_[cur][2].p;
        DArr<P> *im = alias_ ? pos_[cur] : img_[cur];
        a.ix = im[0].p; a.iy = im[1].p; a.iz = im[2].p;
        a.qx = pos_[1 - cur][0].p; a.qy = pos_[1 - cur][1].p; a.qz = pos_[1 - cur][2].p;
        DArr<P> *jm = alias_ ? pos_[1 - cur] : img_[1 - cur];
        a.jx = jm[0].p; a.jy = jm[1].p; a.jz = jm[2].p;
    }

    fnk::Args<T, P> args_granular()
    {
        fnk::Args<T, P> a = args_base();
        set_pos(a, pc_);
        a.fx = frc_[fc_][0].p; a.fy = frc_[fc_][1].p; a.fz = frc_[fc_][2].p;
        a.gx = frc_[1 - fc_][0].p; a.gy = frc_[1 - fc_][1].p; a.gz = frc_[1 - fc_][2].p;
        const auto &kp = kernels_.at("AdvancePosition").args;
        if (kp.count(2)) a.drift = make_box(kp.at(2).s);
        const auto &kc = kernels_.at("ResolveContacts").args;
        if (kc.count(6)) a.force = make_box(kc.at(6).s);
        return a;
    }

    // ---- grid and contact list ---------------------------------------------------------------
    void ensure_grid(const fn_settings *s)
    {
        nb_start_.ensure((size_t)N_ + 1);
        cstart_.ensure((size_t)N_ + 1);
        nb_cnt_.ensure(N_);
        ncnt_.ensure(N_);
        key_.ensure(N_);
        ckey_.ensure(N_);
        rank_.ensure(N_);
        need_.ensure(1);
        for (auto &r : ref_) r.ensure(N_);
        if (!nb_list_.p) grow_list(std::max<long long>(32LL * N_, 1 << 20));
        if (!nb_init_) {
            fnk::fill_int(nb_start_.p, 0, N_ + 1, stream_);
            fnk::fill_int(cstart_.p, 0, N_ + 1, stream_);
            fnk::fill_int(nb_cnt_.p, 0, N_, stream_);
            fnk::fill_int(key_.p, -1, N_, stream_);
            fnk::fill_int(need_.p, 1, 1, stream_);
            nb_init_ = true;
        }
        if (s) {
            const int C = C_of(*s);
            if (C <= 0) throw Error(FN_ERR_ARG, "empty grid");
            if (s->cells_x < 3 || s->cells_y < 3 || s->cells_z < 3)
                throw Error(FN_ERR_UNSUPPORTED, "grid with fewer than three cells along an axis");
            if ((long long)s->cells_x * s->cells_y != s->cells_xy) throw Error(FN_ERR_ARG, "cells_xy mismatch");
            const int nt = s->types_num;
            if (nt <= 0 || (long long)nt * (nt + 1) / 2 > (long long)h_pairs_.size())
                throw Error(FN_ERR_ARG, "pair table smaller than types_num*(types_num+1)/2");
            if (types_checked_ != nt) {   // once per upload of the bead info and types_num
                for (const fn_bead_info &q : h_info_)
                    if (q.type < 0 || q.type >= nt) throw Error(FN_ERR_ARG, "bead type outside [0, types_num)");
                types_checked_ = nt;
            }
            if (C != C_) {
                clear_graphs();
                C_ = C;
                cell_count_.alloc((size_t)C + 1);
                cell_start_.alloc((size_t)C + 1);
                fnk::fill_int(cell_count_.p, 0, C + 1, stream_);
            }
            const int nb = std::max(fnk::scan_blocks_needed(std::max(C_, N_)), 1) + 1;
            if ((int)scan_tmp_.n < nb) {
                clear_graphs();
                scan_tmp_.alloc(nb);
            }
            check_cell_nh(*s);
        }
    }

    // The engine enumerates the standard neighbour stencil; a caller-supplied
    // GridCellNh table must be the one SimulationManager.cpp:2193-2301 builds.
    void check_cell_nh(const fn_settings &s)
    {
        if (cellnh_checked_ || h_cellnh_.empty()) return;
        const int cx = s.cells_x, cy = s.cells_y, cz = s.cells_z;
        if ((int)h_cellnh_.size() != cx * cy * cz) throw Error(FN_ERR_ARG, "GridCellNh table size differs from the grid");
        auto nb = [&](int x, int y, int z) {
            if (s.bc_type == 0) {
                x = x < 0 ? cx - 1 : (x >= cx ? 0 : x);
                y = y < 0 ? cy - 1 : (y >= cy ? 0 : y);
                z = z < 0 ? cz - 1 : (z >= cz ? 0 : z);
            } else if (s.bc_type == 1) {
                if (x < 0 || x >= cx || y < 0 || y >= cy || z < 0 || z >= cz) return -1;
            } else {
                return -1;
            }
            return x + y * cx + z * cx * cy;
        };
        for (int z = 0; z < cz; ++z)
            for (int y = 0; y < cy; ++y)
                for (int x = 0; x < cx; ++x) {
                    const int *t = h_cellnh_[x + y * cx + z * cx * cy].n;
                    int k = 0;
                    bool ok = true;
                    for (int dx = -1; dx < 2; ++dx)
                        for (int dy = -1; dy < 2; ++dy) ok = ok && t[k++] == nb(x + dx, y + dy, z - 1);
                    for (int dy = -1; dy < 2; ++dy) ok = ok && t[k++] == nb(x + 1, y + dy, z);
                    ok = ok && t[k++] == nb(x, y - 1, z);
                    if (!ok) throw Error(FN_ERR_UNSUPPORTED, "GridCellNh table is not the standard half shell");
                }
        cellnh_checked_ = true;
    }

    void grow_list(long long want)
    {
        if (want > (1LL << 31) - 1) throw Error(FN_ERR_CAPACITY, "contact list larger than 2^31 entries");
        if (want <= cap_ && nb_list_.p) return;
        clear_graphs();
        cand_.alloc((size_t)want);
        nb_list_.alloc((size_t)want);
        cap_ = (int)want;
        // entries are only read where a build wrote them, but an overflowing
        // build leaves gaps: keep every entry a valid (bead 0, pair 0) one
        fnk::fill_int(cand_.p, 0, cap_, stream_);
        fnk::fill_int(nb_list_.p, 0, cap_, stream_);
        fnk::fill_int(need_.p, 1, 1, stream_);
    }

    void hash_and_sort(const fnk::Box &b)
    {
        const fn_settings &s = settings_of(K_FILL_GRID);
        ensure_grid(&s);
        ensure_scratch();
        enqueue_sort(pc_, b, nullptr);
        grid_valid_ = true;
    }

    // Counting sort of the beads (positions of buffer pc) into cells.
    void enqueue_sort(int pc, const fnk::Box &b, const int *gate)
    {
        // (the counts are zero: allocated so, and every placement zeroes them)
        fnk::hash(pos_[pc][0].p, pos_[pc][1].p, pos_[pc][2].p, N_, b, ens_net(), boxes_.p, key_.p, ckey_.p, rank_.p,
                  cell_count_.p, gate,
                  stream_);
        fnk::exclusive_scan(cell_count_.p, C_, cell_start_.p, scan_tmp_.p, gate, stream_);
        fnk::place(key_.p, ckey_.p, rank_.p, cell_start_.p, pos_[pc][0].p, pos_[pc][1].p, pos_[pc][2].p, type_.p, tag_.p,
                   loc_.p, N_, sorted_view(), cell_count_.p, C_ + 1, gate, stream_);
    }

    fnk::Sorted sorted_view()
    {
        return fnk::Sorted{so_bead_.p};
    }

    // Largest squared listing radius over the pair types in use: cutoff^2 for
    // lj and st, (max radius + max skin)^2 for hs (kernel.cl:632-659).
    double reach2(const fn_settings &s) const
    {
        double r2 = 0.0;
        for (const fn_pair &p : h_pairs_) {
            if (p.type == 1 || p.type == 3) r2 = std::max(r2, s.cutoff_sq);
            if (p.type == 2) r2 = std::max(r2, (max_radius_ + max_skin_) * (max_radius_ + max_skin_));
        }
        return r2;
    }

    // Candidate radius and stencil for the grid and list settings.  Kept
    // between rebuilds, the candidates are every pair within r = reach + skin
    // (skin a fraction of reach), found over ceil(r / cell) cells per axis;
    // that needs every pair within r seen at its minimum image (always in a
    // finite box, in a periodic one when r < half the smallest side).
    // Rebuilt at every rebuild, they are the pairs within reach in adjacent
    // cells, as the reference lists them.
    void geometry(const fn_settings &g, const fn_settings &l)
    {
        if (g.cells_x != l.cells_x || g.cells_y != l.cells_y || g.cells_z != l.cells_z ||
            g.grid_step.x != l.grid_step.x || g.grid_step.y != l.grid_step.y || g.grid_step.z != l.grid_step.z)
            throw Error(FN_ERR_UNSUPPORTED, "FillGrid and BuildContactList bound to different grids");
        const double r2 = reach2(l);
        const double reach = std::sqrt(r2);
        const bool want = verlet_ == 1 || (verlet_ == 2 && verlet_trial_);
        double skin = want ? reach * skin_permille_ / 1000.0 : 0.0;
        const double lmin = std::min(l.box_size.x, std::min(l.box_size.y, l.box_size.z));
        bool keep = want && skin > 0.0;
        if (l.bc_type == 0 && !(reach + skin < 0.5 * lmin)) keep = false;
        // the filter's cell records pack the three cell coordinates into 31 bits
        const int bx = bits_for(g.cells_x), by = bits_for(g.cells_y), bz = bits_for(g.cells_z);
        if (bx + by + bz > 31) keep = false;
        pack_ = fnk::CellPack{bx, bx + by};
        if (!keep) skin = 0.0;
        const double r = reach + skin;
        fnk::Stencil st{};
        const double step[3] = {g.grid_step.x, g.grid_step.y, g.grid_step.z};
        const int n[3] = {g.cells_x, g.cells_y, g.cells_z};
        for (int k = 0; k < 3; ++k) {
            const double q = (keep && step[k] > 0.0) ? std::ceil(r / step[k] * (1.0 + 1e-12)) : 1.0;
            st.w[k] = (int)std::max(1.0, std::min(q, (double)n[k]));
            st.all[k] = (l.bc_type == 0 && n[k] < 2 * st.w[k] + 1) ? 1 : 0;
        }
        fnk::Stencil ex{};
        for (int k = 0; k < 3; ++k) {
            ex.w[k] = 1;
            ex.all[k] = 0;   // grids have at least three cells per axis
        }
        const double cand_r2 = r * r * (1.0 + 1e-12);
        const double lim = 0.45 * skin;
        if (r2 != reach2_ || cand_r2 != cand_r2_ || keep != keep_cand_ || memcmp(&st, &stencil_, sizeof st) != 0) {
            clear_graphs();
            fnk::fill_int(need_.p, 1, 1, stream_);
        }
        reach2_ = r2;
        cand_r2_ = cand_r2;
        keep_cand_ = keep;
        lim2_ = lim * lim;
        stencil_ = st;
        stencil_exact_ = ex;
    }

    // Bits that hold 0 .. n - 1 (at least one).
    static int bits_for(int n)
    {
        int b = 1;
        while (b < 31 && (1LL << b) < (long long)n) ++b;
        return b;
    }

    // verlet = 2: keeping the candidates pays only while most rebuilds find
    // every bead within skin/2 of its snapshot.  Runs use the one-pass exact
    // rebuild; every 16th run tries the kept candidates, and keeps them for
    // the following runs while fewer than half of the rebuilds had to rebuild
    // them.  The lists are the same either way.
    void tune_verlet(int rebuilds)
    {
        if (verlet_ != 2) return;
        if (!verlet_trial_) {
            if (++verlet_runs_ >= 16) {
                verlet_runs_ = 0;
                verlet_trial_ = true;
            }
            return;
        }
        if (rebuilds < 4) return;
        const int builds = h_stats_[fnk::kStatCandBuilds];   // read back at the end of the run
        if (2 * builds >= rebuilds) verlet_trial_ = false;
    }

    void ensure_scratch()
    {
        so_bead_.ensure(N_);
        cellbead_.ensure(N_);
        if (scratch_.n < (size_t)N_ * K_) {
            clear_graphs();
            scratch_.alloc((size_t)N_ * K_);
        }
    }

    // Search, offsets, compaction.  final_list: the pairs BuildContactList
    // lists (exact rule, reference stencil) into the contact list itself;
    // otherwise the candidates within reach + skin into the candidate list.
    // Gated by `gate`.  sync (per-kernel path, ungated): the host checks the
    // row and list capacities before compacting and grows them as needed.
    void search_and_compact(const fnk::Box &b, bool final_list, bool sync, const int *gate,
                            const fnk::Snapshot &snap = fnk::Snapshot{})
    {
        for (;;) {
            if (sync) {
                fnk::fill_int(stats_.p + fnk::kStatListOver, 0, 1, stream_);
                fnk::fill_int(stats_.p + fnk::kStatDense, 0, 1, stream_);
            }
            // every bead has a slot (k_place clamps the ones outside the grid
            // into it), so the search writes every bead's count
            fnk::search(cell_start_.p, C_, sorted_view(), pairs_.p, rad_d_.p, skin_d_.p, b, ens_net(),
                        boxes_.p, final_list ? stencil_exact_ : stencil_, final_list ? reach2_ : cand_r2_, ens_grids3_,
                        final_list,
                        ncnt_.p, scratch_.p, K_, stats_.p, N_, staged_, flat_, gate, stream_);
            // offsets: the contact list's (exact) or the candidates' CSR
            int *starts = final_list ? nb_start_.p : cstart_.p;
            fnk::exclusive_scan(ncnt_.p, N_, starts, scan_tmp_.p, gate, stream_);
            if (sync) {
                int tot = 0, over = 0, dense = 0;
                HIPCK(hipMemcpyAsync(&tot, starts + N_, sizeof(int), hipMemcpyDeviceToHost, stream_));
                HIPCK(hipMemcpyAsync(&over, stats_.p + fnk::kStatListOver, sizeof(int), hipMemcpyDeviceToHost, stream_));
                HIPCK(hipMemcpyAsync(&dense, stats_.p + fnk::kStatDense, sizeof(int), hipMemcpyDeviceToHost, stream_));
                HIPCK(hipStreamSynchronize(stream_));
                if (dense > 0)
                    throw Error(FN_ERR_UNSUPPORTED, fmt("%d beads lay in grid cells holding more than %d beads (the "
                                                     "reference keeps 64 per cell)", dense, fnk::kMaxCellBeads));
                if (over > 0) {
                    grow_rows();
                    continue;
                }
                if (tot > cap_) grow_list((long long)tot * 2);
                last_total_ = tot;
            }
            fnk::compact(ncnt_.p, starts, scratch_.p, K_, N_, final_list ? nb_list_.p : cand_.p, cap_,
                         final_list ? nb_cnt_.p : nullptr, stats_.p, snap, gate, stream_);
            return;
        }
    }

    // The listing rule on the candidates at the positions of buffer pc.
    void enqueue_filter(int pc, const fnk::Box &bg, const fnk::Box &bl)
    {
        fnk::FilterArgs f{};
        f.cb = cellbead_.p;
        f.px = pos_[pc][0].p; f.py = pos_[pc][1].p; f.pz = pos_[pc][2].p;
        f.pack = pack_;
        // a bead inside the grid has |coordinate| < cells * step on every
        // axis; rounding two of them to fp32 moves their difference by at
        // most 2 * 2^-24 of that
        // (an ensemble: the bound of its largest grid, the lowest guard)
        double extent = 0.0;
        for (const fnk::Box &eb : ens_boxes_.empty() ? std::vector<fnk::Box>{bg} : ens_boxes_)
            extent = std::max(extent, std::max(eb.cx * eb.sx, std::max(eb.cy * eb.sy, eb.cz * eb.sz)));
        f.delta = std::ldexp(extent, -23) * (1.0 + 1e-9) + 1e-300;
        f.cutoff = std::sqrt(bl.cutoff_sq);
        f.hbs_guard = 1e300;
        for (const fnk::Box &eb : ens_boxes_.empty() ? std::vector<fnk::Box>{bl} : ens_boxes_) {
            const double rh = std::sqrt(eb.hbs);
            f.hbs_guard = std::min(f.hbs_guard, eb.hbs - (12.0 * f.delta * (2.0 * rh + f.delta) + 4e-14 * eb.hbs));
        }
        f.n = N_;
        f.grid = bg;
        f.list = bl;
        f.cand_start = cstart_.p; f.cand = cand_.p; f.cap = cap_;
        f.list_start = nb_start_.p;
        f.pairs = pairs_.p;
        f.n_pairs = (int)h_pairs_.size();
        if (!ens_boxes_.empty()) {        // entries are local to their network's table
            f.net = ens_net();
            f.boxes = boxes_.p;
            f.n_pairs = 0;
            for (const fnk::Box &eb : ens_boxes_) f.n_pairs = std::max(f.n_pairs, eb.types * (eb.types + 1) / 2);
        }
        f.radius = rad_d_.p; f.skin = skin_d_.p;
        f.list_out = nb_list_.p; f.cnt = nb_cnt_.p; f.stats = stats_.p;
        f.need = need_.p;
        fnk::filter(f, stream_);
    }

    // One rebuild of the fast path at the positions of buffer pc.  With kept
    // candidates: the displacement check, the gated candidate build with its
    // snapshot, then the filter.  Otherwise the reference's rebuild in one
    // pass: sort, exact search, compaction into the contact list (and the
    // candidates are marked stale).
    void enqueue_rebuild(int pc, const fnk::Box &bg, const fnk::Box &bl)
    {
        const double *p[3] = {pos_[pc][0].p, pos_[pc][1].p, pos_[pc][2].p};
        if (!keep_cand_) {
            // (kept candidates are marked stale by geometry() when they are
            // switched on, so this path leaves need_ alone)
            enqueue_sort(pc, bg, nullptr);
            search_and_compact(bl, true, false, nullptr);
            return;
        }
        fnk::disp_check(p[0], p[1], p[2], ref_[0].p, ref_[1].p, ref_[2].p, N_, bl, bg, ens_net(), boxes_.p, pack_,
                        lim2_, need_.p, cellbead_.p, key_.p, stream_);
        // gated by the flag: the build, whose compaction also records the
        // positions it was made at; the filter clears the flag
        enqueue_sort(pc, bg, need_.p);
        search_and_compact(bl, false, false, need_.p, fnk::Snapshot{p[0], p[1], p[2], ref_[0].p, ref_[1].p, ref_[2].p});
        enqueue_filter(pc, bg, bl);
    }

    void grow_rows()
    {
        set_rows(2 * K_);
        ensure_scratch();
    }

    // Search rows of k entries: the rows of all beads are one array whose
    // offsets are ints, so N * k stays below 2^31.
    void set_rows(long long k)
    {
        if ((long long)N_ * k > (1LL << 31) - 1)
            throw Error(FN_ERR_CAPACITY, fmt("search rows of %lld entries for %d beads exceed 2^31 entries", k, N_));
        K_ = (int)k;
    }

    // BuildContactList of the per-kernel path, after FillGrid's sort: the
    // exact search into the contact list, sizes checked on the host.
    void build_list(const fnk::Box &bl)
    {
        ensure_grid(&settings_of(K_BUILD_LIST));
        ensure_scratch();
        geometry(settings_of(K_FILL_GRID), settings_of(K_BUILD_LIST));
        fnk::set_int(need_.p, 1, nullptr, stream_);
        search_and_compact(bl, true, true, nullptr);
    }

    void check_capacity()
    {
        int st[fnk::kStatCount];
        HIPCK(hipMemcpy(st, stats_.p, sizeof st, hipMemcpyDeviceToHost));
        check_capacity(st);
    }
    // The same from statistics already read back (st: kStatCount values).
    void check_capacity(const int *st)
    {
        const int tot = st[fnk::kStatCandMax], over = st[fnk::kStatListOver];
        if (st[fnk::kStatDense] > 0)
            throw Error(FN_ERR_UNSUPPORTED, fmt("%d beads lay in grid cells holding more than %d beads (the reference keeps "
                                             "64 per cell); the run is invalid", st[fnk::kStatDense], fnk::kMaxCellBeads));
        last_total_ = std::max(last_total_, tot);
        // the longest rows the searches needed (counted past the row length)
        rows_exact_ = std::max(rows_exact_, st[fnk::kStatMaxList]);
        rows_cand_ = std::max(rows_cand_, st[fnk::kStatMaxCand]);
        if (over > 0) {
            const int need = std::max(st[fnk::kStatMaxList], st[fnk::kStatMaxCand]), was = K_;
            long long k = 2LL * K_;
            while (k < 2LL * need) k *= 2;
            set_rows(k);
            ensure_scratch();
            throw Error(FN_ERR_CAPACITY, fmt("a contact list needed %d entries of a %d-entry search row; the run "
                                             "is invalid and the row length is now %d", need, was, K_));
        }
        if (tot > cap_) {
            grow_list((long long)tot * 2);
            throw Error(FN_ERR_CAPACITY, fmt("contact list needed %d entries, %d were reserved; the run is invalid "
                                             "and the capacity has been raised", tot, cap_));
        }
    }

    // ---- fused chunks -----------------------------------------------------------------------------
    int chunk_len(int guf) const
    {
        const int target = chunk_ < 1 ? 1 : chunk_;
        return guf * std::max(1, target / guf);
    }

    // Called after each step's bead kernel with the step's index in the chunk,
    // the position buffer the step read and the force buffer it wrote.
    using StepHook = std::function<void(int, int, int)>;

    void enqueue_chunk(int len, int guf, bool spr, bool ang, bool first, bool last, Events *ev,
                       const StepHook *hook = nullptr)
    {
        const fnk::Box bd = kbox(K_ADVANCE_POSITION);
        const fnk::Box bf = kbox(K_FILL_GRID);
        const fnk::Box bl = kbox(K_BUILD_LIST);
        const fnk::Box bc = kbox(K_RESOLVE_CONTACTS);
        int pc = pc_, fc = fc_;
        const bool chain = use_chain();
        if (chain && cmp_ok_)        // the chain kernel's status copies, whatever ran before
            fnk::chain_status(ch_spring_.p, ch_angle_.p, sp_status_.p, an_status_.p, bond_code_.p, N_, stream_);
        if (first) {
            fnk::Args<T, P> a = args_base();
            a.drift = bd;
            set_pos(a, pc);
            a.fx = frc_[fc][0].p; a.fy = frc_[fc][1].p; a.fz = frc_[fc][2].p;
            fnk::Launch<T, P>::drift_first(a, stream_);
            pc ^= 1;
        }
        for (int k = 0; k < len; ++k) {
            if (k % guf == 0) {
                mark(ev ? &ev->rebuild : nullptr);
                enqueue_rebuild(pc, bf, bl);
                mark(ev ? &ev->rebuild : nullptr);
            }
            const bool is_last = last && k == len - 1;
            fnk::Args<T, P> a = args_base();
            a.drift = bd;
            a.force = bc;
            a.n_springs = n_spr_;
            a.n_angles = n_ang_;
            a.step_off = k;
            set_pos(a, pc);
            a.fx = frc_[fc][0].p; a.fy = frc_[fc][1].p; a.fz = frc_[fc][2].p;
            a.gx = frc_[1 - fc][0].p; a.gy = frc_[1 - fc][1].p; a.gz = frc_[1 - fc][2].p;
            std::vector<hipEvent_t> *sv = ev ? &ev->step : nullptr;
            mark(sv);
            if (spr && !chain) fnk::Launch<T, P>::spring_terms(a, S_, stream_);
            mark(sv);
            if (ang && !chain) fnk::Launch<T, P>::angle_terms(a, A_, stream_);
            mark(sv);
            fnk::Launch<T, P>::bead_step(a, spr, ang, !is_last, is_last, chain, block_, stream_);
            mark(sv);
            if (hook) (*hook)(k, pc, 1 - fc);
            fc ^= 1;
            if (!is_last) pc ^= 1;
        }
        fnk::add_int(step_.p, len, stream_);
    }

    bool use_chain() const { return chain_on_ && chain_ok_ && n_spr_ == S_ && n_ang_ == A_; }

    // Settings of a kernel as a Box; in an ensemble every kernel reads its
    // network's Box and this one (the first network's) only fills the slot.
    fnk::Box kbox(int kid)
    {
        if (!ens_boxes_.empty()) return ens_boxes_[0];
        return make_box(settings_of(kid));
    }

    // ---- ensemble ---------------------------------------------------------------------------
    // Independent networks in one engine: beads, bonds and pair tables
    // concatenated (bond indices shifted), every bead tagged with its network,
    // every network with its own Box (box, grid, integrator constants) whose
    // cells are numbered after the previous networks' and whose pair entries
    // start at pair_base.  Contacts never cross networks: a bead only searches
    // its own network's cells.
    // The networks' records concatenated (bond indices shifted) with their
    // Boxes, after the checks of fn_batch_create.
    struct Packed {
        std::vector<fn_bead> beads;
        std::vector<fn_bead_info> info;
        std::vector<fn_spring> sp;
        std::vector<fn_angle> an;
        std::vector<fn_pair> pr;
        std::vector<fnk::Box> boxes;
        std::vector<int> net;
        std::vector<fn_settings> sets;
        std::vector<int> b0{0}, s0{0}, a0{0};
        long long cells = 0;
        bool finite = true;
    };

    Packed pack_ensemble(const fn_network *nets, int n)
    {
        if (n <= 0 || !nets) throw Error(FN_ERR_ARG, "an ensemble needs at least one network");
        Packed pk;
        auto &beads = pk.beads;
        auto &info = pk.info;
        auto &sp = pk.sp;
        auto &an = pk.an;
        auto &pr = pk.pr;
        auto &boxes = pk.boxes;
        auto &net = pk.net;
        auto &sets = pk.sets;
        auto &b0 = pk.b0;
        auto &s0 = pk.s0;
        auto &a0 = pk.a0;
        long long &cells = pk.cells;
        bool &finite = pk.finite;
        for (int k = 0; k < n; ++k) {
            const fn_network &w = nets[k];
            if (w.n_beads <= 0 || !w.beads || !w.info || !w.settings || w.n_pairs <= 0 || !w.pairs)
                throw Error(FN_ERR_ARG, fmt("network %d: beads, bead info, pair table and settings are required", k));
            if (w.n_springs < 0 || w.n_angles < 0 || (w.n_springs && !w.springs) || (w.n_angles && !w.angles))
                throw Error(FN_ERR_ARG, fmt("network %d: bond counts and records disagree", k));
            const fn_settings &st = *w.settings;
            const int nt = st.types_num;
            if (st.cells_x < 3 || st.cells_y < 3 || st.cells_z < 3 || (long long)st.cells_x * st.cells_y != st.cells_xy)
                throw Error(FN_ERR_UNSUPPORTED, fmt("network %d: grid with fewer than three cells along an axis", k));
            if (nt <= 0 || (long long)nt * (nt + 1) / 2 > w.n_pairs || w.n_pairs > (1 << (31 - fnk::kPidShift)))
                throw Error(FN_ERR_ARG, fmt("network %d: pair table does not match types_num", k));
            const int bb = (int)beads.size(), sb = (int)sp.size(), pb = (int)pr.size();
            for (int i = 0; i < w.n_beads; ++i) {
                if (w.info[i].type < 0 || w.info[i].type >= nt)
                    throw Error(FN_ERR_ARG, fmt("network %d: bead type outside [0, types_num)", k));
                beads.push_back(w.beads[i]);
                info.push_back(w.info[i]);
                net.push_back(k);
            }
            for (int q = 0; q < w.n_springs; ++q) {
                fn_spring r = w.springs[q];
                if (r.p1 < 0 || r.p1 >= w.n_beads || r.p2 < 0 || r.p2 >= w.n_beads)
                    throw Error(FN_ERR_ARG, fmt("network %d: spring bead index out of range", k));
                r.p1 += bb;
                r.p2 += bb;
                sp.push_back(r);
            }
            for (int q = 0; q < w.n_angles; ++q) {
                fn_angle r = w.angles[q];
                if (r.b1 < 0 || r.b1 >= w.n_beads || r.b2 < 0 || r.b2 >= w.n_beads || r.b3 < 0 || r.b3 >= w.n_beads)
                    throw Error(FN_ERR_ARG, fmt("network %d: angle bead index out of range", k));
                if (r.type == 1 && (r.s1 < 0 || r.s1 >= w.n_springs || r.s2 < 0 || r.s2 >= w.n_springs))
                    throw Error(FN_ERR_ARG, fmt("network %d: HR_FL angle spring index out of range", k));
                r.b1 += bb;
                r.b2 += bb;
                r.b3 += bb;
                r.s1 += sb;
                r.s2 += sb;
                an.push_back(r);
            }
            for (int q = 0; q < w.n_pairs; ++q) pr.push_back(w.pairs[q]);
            fnk::Box bx = make_box(st);
            bx.cell_base = (int)cells;
            bx.pair_base = pb;
            cells += (long long)st.cells_xy * st.cells_z;
            if (cells >= (1LL << 31) - 1) throw Error(FN_ERR_CAPACITY, "more than 2^31 grid cells in the ensemble");
            boxes.push_back(bx);
            sets.push_back(st);
            finite = finite && st.bc_type == 1;
            b0.push_back((int)beads.size());
            s0.push_back((int)sp.size());
            a0.push_back((int)an.size());
        }
        return pk;
    }

    void load_ensemble(const fn_network *nets, int n) override
    {
        adj_.valid = false;
        if (!bufs_.empty()) throw Error(FN_ERR_ARG, "an ensemble is loaded into an engine without buffers");
        Packed pk = pack_ensemble(nets, n);
        create_buffer(FN_BUF_BEAD, sizeof(fn_bead), (int)pk.beads.size(), pk.beads.data());
        create_buffer(FN_BUF_BEAD_INFO, sizeof(fn_bead_info), (int)pk.info.size(), pk.info.data());
        if (!pk.sp.empty()) create_buffer(FN_BUF_SPRING, sizeof(fn_spring), (int)pk.sp.size(), pk.sp.data());
        if (!pk.an.empty()) create_buffer(FN_BUF_ANGLE, sizeof(fn_angle), (int)pk.an.size(), pk.an.data());
        create_buffer(FN_BUF_PAIR, sizeof(fn_pair), (int)pk.pr.size(), pk.pr.data());
        adopt_ensemble(pk);
        cell_count_.alloc((size_t)C_ + 1);
        cell_start_.alloc((size_t)C_ + 1);
        fnk::fill_int(cell_count_.p, 0, C_ + 1, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        clear_graphs();
    }

    // New records for every network of a loaded ensemble, of the same sizes
    // (beads, bonds, pair entries, grid cells); runs continue from them.
    void reload_ensemble(const fn_network *nets, int n) override
    {
        adj_.valid = false;
        if (ens_boxes_.empty()) throw Error(FN_ERR_ARG, "no ensemble is loaded");
        Packed pk = pack_ensemble(nets, n);
        if (n != (int)ens_boxes_.size() || pk.b0 != ens_b0_ || pk.s0 != ens_s0_ || pk.a0 != ens_a0_ ||
            pk.pr.size() != h_pairs_.size() || pk.cells != C_)
            throw Error(FN_ERR_MISMATCH, "the networks differ in size from the loaded ones");
        upload(FN_BUF_BEAD, (int)pk.beads.size(), pk.beads.data());
        upload(FN_BUF_BEAD_INFO, (int)pk.info.size(), pk.info.data());
        if (!pk.sp.empty()) upload(FN_BUF_SPRING, (int)pk.sp.size(), pk.sp.data());
        if (!pk.an.empty()) upload(FN_BUF_ANGLE, (int)pk.an.size(), pk.an.data());
        upload(FN_BUF_PAIR, (int)pk.pr.size(), pk.pr.data());
        adopt_ensemble(pk);
        HIPCK(hipStreamSynchronize(stream_));
        clear_graphs();
    }

    void adopt_ensemble(const Packed &pk)
    {
        ens_boxes_ = pk.boxes;
        ens_sets_ = pk.sets;
        ens_b0_ = pk.b0;
        ens_s0_ = pk.s0;
        ens_a0_ = pk.a0;
        ens_finite_ = pk.finite;
        boxes_.put(pk.boxes, stream_);
        net_.put(pk.net, stream_);
        C_ = (int)pk.cells;
        ens_grids3_ = true;
        for (const fnk::Box &eb : ens_boxes_) ens_grids3_ = ens_grids3_ && eb.cx >= 3 && eb.cy >= 3 && eb.cz >= 3;
    }

    void prepare_ensemble_run(bool spr, bool ang)
    {
        (void)spr;
        (void)ang;
        prepare();
        n_spr_ = S_;
        n_ang_ = A_;
        check_alias(ens_finite_ ? 1 : 0);
        ensure_grid(nullptr);
        const int nb = std::max(fnk::scan_blocks_needed(std::max(C_, N_)), 1) + 1;
        if ((int)scan_tmp_.n < nb) {
            clear_graphs();
            scan_tmp_.alloc(nb);
        }
        ensure_scratch();
        // one rebuild path (the exact search over every network's own 3x3x3
        // stencil); the prefilter distance covers every network's reach
        double r2 = 0.0, cut2 = 0.0;
        for (const fn_settings &st : ens_sets_) cut2 = std::max(cut2, st.cutoff_sq);
        fn_settings any{};
        any.cutoff_sq = cut2;
        r2 = reach2(any);
        fnk::Stencil ex{};
        for (int k = 0; k < 3; ++k) ex.w[k] = 1;
        // kept candidates as geometry() decides them for one network, with
        // one skin and candidate reach for all networks (from the largest
        // listing reach) and a stencil wide enough for the finest grid; the
        // search visits a periodic grid narrower than it whole
        const double reach = std::sqrt(r2);
        const bool want = verlet_ == 1 || (verlet_ == 2 && verlet_trial_);
        double skin = want ? reach * skin_permille_ / 1000.0 : 0.0;
        bool keep = want && skin > 0.0;
        int mx = 1, my = 1, mz = 1;
        for (const fnk::Box &eb : ens_boxes_) {
            const double lmin = std::min(eb.lx, std::min(eb.ly, eb.lz));
            if (eb.bc == 0 && !(reach + skin < 0.5 * lmin)) keep = false;
            mx = std::max(mx, eb.cx);
            my = std::max(my, eb.cy);
            mz = std::max(mz, eb.cz);
        }
        const int bx = bits_for(mx), by = bits_for(my), bz = bits_for(mz);
        if (bx + by + bz > 31) keep = false;
        pack_ = fnk::CellPack{bx, bx + by};
        if (!keep) skin = 0.0;
        const double r = reach + skin;
        fnk::Stencil st{};
        for (int k = 0; k < 3; ++k) {
            int w = 1;
            for (const fnk::Box &eb : ens_boxes_) {
                const double step = k == 0 ? eb.sx : (k == 1 ? eb.sy : eb.sz);
                if (keep && step > 0.0) w = std::max(w, (int)std::ceil(r / step * (1.0 + 1e-12)));
            }
            st.w[k] = w;
            st.all[k] = 0;
        }
        const double cand_r2 = r * r * (1.0 + 1e-12);
        if (r2 != reach2_ || cand_r2 != cand_r2_ || keep != keep_cand_ || memcmp(&st, &stencil_, sizeof st) != 0 ||
            cand_dirty_) {
            clear_graphs();
            fnk::fill_int(need_.p, 1, 1, stream_);
            cand_dirty_ = false;
        }
        reach2_ = r2;
        cand_r2_ = cand_r2;
        keep_cand_ = keep;
        lim2_ = (0.45 * skin) * (0.45 * skin);
        stencil_ = st;
        stencil_exact_ = ex;
    }

    const int *ens_net() const { return ens_boxes_.empty() ? nullptr : net_.p; }

    // Records of one network, with its bond indices shifted back.
    void ensemble_read(int k, int kind, void *out) override
    {
        if (ens_boxes_.empty()) throw Error(FN_ERR_ARG, "the engine holds no ensemble");
        if (k < 0 || k >= (int)ens_boxes_.size()) throw Error(FN_ERR_ARG, fmt("no network %d", k));
        const int b0 = ens_b0_[k], b1 = ens_b0_[k + 1], s0 = ens_s0_[k], s1 = ens_s0_[k + 1];
        const int a0 = ens_a0_[k], a1 = ens_a0_[k + 1];
        switch (kind) {
        case FN_BUF_BEAD: {
            std::vector<fn_bead> all((size_t)pad64(N_));
            download_beads(all.data());
            memcpy(out, all.data() + b0, sizeof(fn_bead) * (size_t)(b1 - b0));
            break;
        }
        case FN_BUF_SPRING: {
            std::vector<int> st;
            sp_status_.get(st, S_, stream_);
            HIPCK(hipStreamSynchronize(stream_));
            fn_spring *o = (fn_spring *)out;
            for (int q = s0; q < s1; ++q) {
                fn_spring r = h_springs_[q];
                r.status = st[q];
                r.p1 -= b0;
                r.p2 -= b0;
                o[q - s0] = r;
            }
            break;
        }
        case FN_BUF_ANGLE: {
            std::vector<int> st;
            an_status_.get(st, A_, stream_);
            HIPCK(hipStreamSynchronize(stream_));
            fn_angle *o = (fn_angle *)out;
            for (int q = a0; q < a1; ++q) {
                fn_angle r = h_angles_[q];
                r.status = st[q];
                r.b1 -= b0;
                r.b2 -= b0;
                r.b3 -= b0;
                r.s1 -= s0;
                r.s2 -= s0;
                o[q - a0] = r;
            }
            break;
        }
        default: throw Error(FN_ERR_ARG, "ensemble records are beads, springs or angles");
        }
    }

    // ---- autotuner --------------------------------------------------------------------------
    // Runs n steps exactly as run_steps does while trying the settings that
    // leave every result bitwise the same (chunk length, chain kernel,
    // staged search, kept candidates), one knob at a time: each value runs
    // one chunk (graph capture, warm caches) and one timed chunk, and the
    // fastest value per step is kept before the next knob.  Steps left over
    // run with the chosen settings, which stay set.
    void autotune(int n, int guf, bool spr, bool ang, std::vector<fn_tune_trial> &trials) override
    {
        adj_.valid = false;
        trials.clear();
        prep_fast(n, guf, spr, ang);
        if ((long long)last_total_ * 3 / 2 + 1024 > (long long)cap_) grow_list((long long)last_total_ * 2 + 1024);
        fnk::fill_int(stats_.p, 0, fnk::kStatCount, stream_);
        const bool ens = !ens_boxes_.empty();
        std::vector<std::pair<int, std::vector<int>>> knobs;
        std::vector<int> chunks;
        for (int c : {32, 64, 128, 256}) {
            const int l = guf * std::max(1, c / guf);
            if (std::find(chunks.begin(), chunks.end(), l) == chunks.end()) chunks.push_back(l);
        }
        if (chain_ok_ && n_spr_ == S_ && n_ang_ == A_) knobs.push_back({FN_TUNE_CHAIN, {1, 0}});
        if (!ens) knobs.push_back({FN_TUNE_STAGED, {0, 1}});
        knobs.push_back({FN_TUNE_VERLET, {verlet_, 0, 1}});
        // the kept candidates' skin (used while verlet keeps them): a thinner
        // skin filters fewer candidates and rebuilds them more often
        knobs.push_back({FN_TUNE_SKIN, {skin_permille_, 300, 150, 75}});
        knobs.push_back({FN_TUNE_BLOCK, {block_, 256, 128, 512}});
        knobs.push_back({FN_TUNE_CHUNK, chunks});       // last: its trials are the longest
        static const char *names[] = {"chunk", "chain", "staged", "verlet", "block", "skin"};
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPCK(hipEventCreate(&e0));
        HIPCK(hipEventCreate(&e1));
        int done = 0, rb = 0;
        auto chunk = [&](bool timed) {
            const int len = std::min(chunk_len(guf), n - done);
            if (timed) HIPCK(hipEventRecord(e0, stream_));
            run_chunk(len, guf, spr, ang, done == 0, done + len == n);
            rb += (len + guf - 1) / guf;
            done += len;
            float ms = 0.0f;
            if (timed) {
                HIPCK(hipEventRecord(e1, stream_));
                HIPCK(hipEventSynchronize(e1));
                HIPCK(hipEventElapsedTime(&ms, e0, e1));
            }
            return (double)ms / len;
        };
        auto apply = [&](int knob, int v) {
            set_option(names[knob], v);
            if (ens) prepare_ensemble_run(spr, ang);
            else geometry(settings_of(K_FILL_GRID), settings_of(K_BUILD_LIST));
        };
        try {
            for (auto &kv : knobs) {
                const int knob = kv.first;
                int best_v = get_option(names[knob]);
                double best = 1e300;
                std::vector<int> tried;
                for (int v : kv.second) {
                    if (std::find(tried.begin(), tried.end(), v) != tried.end()) continue;
                    tried.push_back(v);
                    if (n - done < 2 * guf) break;
                    apply(knob, v);
                    chunk(false);
                    if (done >= n) break;
                    const double t = chunk(true);
                    trials.push_back(fn_tune_trial{knob, v, t});
                    log(fmt("autotune: %s = %d: %.4f ms per step", names[knob], v, t));
                    if (t < best) {
                        best = t;
                        best_v = v;
                    }
                }
                if (get_option(names[knob]) != best_v) apply(knob, best_v);
            }
            while (done < n) chunk(false);
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        HIPCK(hipEventDestroy(e0));
        HIPCK(hipEventDestroy(e1));
        steps_ += n;
        rebuilds_ += rb;
        HIPCK(hipStreamSynchronize(stream_));
        collect_timing();
        check_capacity();
    }

    // ---- adjoint ----------------------------------------------------------------------------
    struct Part {
        void *p;
        size_t bytes;
    };
    static size_t padded(size_t b) { return (b + 255) / 256 * 256; }

    // The device state a run continues from: the positions and images the
    // next step reads (buffer pc), velocity_old, the previous force (buffer
    // fc), bond statuses and break steps, the step counter; with `full` also
    // the read-back velocity and force_old and what the exports and stats
    // read (cells, contact list, counters).
    std::vector<Part> state_parts(int pc, int fc, bool full)
    {
        std::vector<Part> v;
        for (int q = 0; q < 3; ++q) v.push_back({pos_[pc][q].p, sizeof(P) * N_});
        if (!alias_)
            for (int q = 0; q < 3; ++q) v.push_back({img_[pc][q].p, sizeof(P) * N_});
        for (int q = 0; q < 3; ++q) v.push_back({vold_[q].p, sizeof(T) * N_});
        for (int q = 0; q < 3; ++q) v.push_back({frc_[fc][q].p, sizeof(T) * N_});
        if (S_) {
            v.push_back({sp_status_.p, sizeof(int) * S_});
            v.push_back({sp_break_.p, sizeof(int) * S_});
        }
        if (A_) v.push_back({an_status_.p, sizeof(int) * A_});
        v.push_back({step_.p, sizeof(int)});
        if (full) {
            for (int q = 0; q < 3; ++q) v.push_back({vel_[q].p, sizeof(T) * N_});
            for (int q = 0; q < 3; ++q) v.push_back({frc_[1 - fc][q].p, sizeof(T) * N_});   // read back as force_old
            v.push_back({key_.p, sizeof(int) * N_});
            v.push_back({nb_start_.p, sizeof(int) * ((size_t)N_ + 1)});
            v.push_back({nb_cnt_.p, sizeof(int) * N_});
            v.push_back({nb_list_.p, sizeof(int) * (size_t)cap_});
            v.push_back({stats_.p, sizeof(int) * fnk::kStatCount});
        }
        return v;
    }
    static size_t parts_bytes(const std::vector<Part> &v)
    {
        size_t b = 0;
        for (const Part &x : v) b += padded(x.bytes);
        return b;
    }
    void save_parts(const std::vector<Part> &v, unsigned char *dst)
    {
        for (const Part &x : v) {
            HIPCK(hipMemcpyAsync(dst, x.p, x.bytes, hipMemcpyDeviceToDevice, stream_));
            dst += padded(x.bytes);
        }
    }
    void load_parts(const std::vector<Part> &v, const unsigned char *src)
    {
        for (const Part &x : v) {
            HIPCK(hipMemcpyAsync(x.p, src, x.bytes, hipMemcpyDeviceToDevice, stream_));
            src += padded(x.bytes);
        }
    }

    // ---- checkpoints ----------------------------------------------------------------------
    // File layout: the header, then the bead records (fp64 AoS, which holds
    // the fp32 build's values exactly), the spring statuses and break steps,
    // and the angle statuses.
    struct CkHeader {
        char magic[8];
        int precision, n_beads, n_springs, n_angles;
        int step_counter, pad;
        long long steps, rebuilds;
    };
    static constexpr char kCkMagic[8] = {'F', 'N', 'C', 'K', 'P', 'T', '0', '1'};

    void checkpoint_save(const char *path) override
    {
        if (!path) throw Error(FN_ERR_ARG, "checkpoint path is null");
        if (N_ <= 0) throw Error(FN_ERR_ARG, "no bead buffer");
        HIPCK(hipStreamSynchronize(stream_));
        CkHeader h{};
        memcpy(h.magic, kCkMagic, sizeof h.magic);
        h.precision = precision;
        h.n_beads = N_;
        h.n_springs = S_;
        h.n_angles = A_;
        h.steps = steps_;
        h.rebuilds = rebuilds_;
        std::vector<fn_bead> b(N_);
        download_beads(b.data());
        std::vector<int> sst, sbr, ast, stp;
        sp_status_.get(sst, S_, stream_);
        sp_break_.get(sbr, S_, stream_);
        an_status_.get(ast, A_, stream_);
        step_.get(stp, 1, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        h.step_counter = stp[0];
        FILE *f = fopen(path, "wb");
        if (!f) throw Error(FN_ERR_ARG, fmt("unable to open %s for writing", path));
        bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(b.data(), sizeof(fn_bead), N_, f) == (size_t)N_;
        if (S_) ok = ok && fwrite(sst.data(), sizeof(int), S_, f) == (size_t)S_ &&
                     fwrite(sbr.data(), sizeof(int), S_, f) == (size_t)S_;
        if (A_) ok = ok && fwrite(ast.data(), sizeof(int), A_, f) == (size_t)A_;
        ok = (fclose(f) == 0) && ok;
        if (!ok) throw Error(FN_ERR_ARG, fmt("unable to write the checkpoint %s", path));
    }

    void checkpoint_load(const char *path) override
    {
        if (!path) throw Error(FN_ERR_ARG, "checkpoint path is null");
        FILE *f = fopen(path, "rb");
        if (!f) throw Error(FN_ERR_ARG, fmt("unable to open %s", path));
        std::unique_ptr<FILE, int (*)(FILE *)> guard(f, fclose);
        CkHeader h{};
        if (fread(&h, sizeof h, 1, f) != 1 || memcmp(h.magic, kCkMagic, sizeof h.magic) != 0)
            throw Error(FN_ERR_MISMATCH, fmt("%s is not a checkpoint of this library", path));
        if (h.precision != precision || h.n_beads != N_ || h.n_springs != S_ || h.n_angles != A_)
            throw Error(FN_ERR_MISMATCH, fmt("checkpoint of precision %d with %d beads, %d springs, %d angles; the "
                                             "engine has precision %d with %d, %d, %d", h.precision, h.n_beads,
                                             h.n_springs, h.n_angles, precision, N_, S_, A_));
        std::vector<fn_bead> b(N_);
        std::vector<int> sst(S_), sbr(S_), ast(A_);
        bool ok = fread(b.data(), sizeof(fn_bead), N_, f) == (size_t)N_;
        if (S_) ok = ok && fread(sst.data(), sizeof(int), S_, f) == (size_t)S_ &&
                     fread(sbr.data(), sizeof(int), S_, f) == (size_t)S_;
        if (A_) ok = ok && fread(ast.data(), sizeof(int), A_, f) == (size_t)A_;
        if (!ok) throw Error(FN_ERR_MISMATCH, fmt("%s is truncated", path));
        adj_.valid = false;
        HIPCK(hipStreamSynchronize(stream_));
        upload_beads(b.data());
        if (S_) {
            sp_status_.put(sst, stream_);
            sp_break_.put(sbr, stream_);
        }
        if (A_) an_status_.put(ast, stream_);
        std::vector<int> stp(1, h.step_counter);
        step_.put(stp, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        steps_ = h.steps;
        rebuilds_ = h.rebuilds;
        grid_valid_ = false;
    }

    // ---- observables ----------------------------------------------------------------------
    void group_force(const int *group, double *out) override
    {
        if (!group || !out) throw Error(FN_ERR_ARG, "group and output are required");
        if (N_ <= 0) throw Error(FN_ERR_ARG, "no bead buffer");
        const int nn = n_networks();
        std::vector<int> gi, goff(nn + 1, 0);
        for (int i = 0; i < N_; ++i)
            if (group[i] != 0) gi.push_back(i);
        int max_len = 0;
        for (int k = 0; k < nn; ++k) {
            const int b1 = ens_boxes_.empty() ? N_ : ens_b0_[k + 1];
            goff[k + 1] = (int)(std::lower_bound(gi.begin(), gi.end(), b1) - gi.begin());
            max_len = std::max(max_len, goff[k + 1] - goff[k]);
        }
        rf_idx_.put(gi, stream_);
        rf_off_.put(goff, stream_);
        rf_part_.ensure(fnk::kSumBlocks * (size_t)nn);
        rf_out_.ensure(nn);
        fnk::group_force(frc_[fc_][0].p, rf_idx_.p, rf_off_.p, nn, max_len, rf_part_.p, rf_out_.p, stream_);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(out, rf_out_.p, sizeof(double) * nn, hipMemcpyDeviceToHost, stream_));
        HIPCK(hipStreamSynchronize(stream_));
    }

    // fn_group_force with every device buffer in the fixed workspace
    // (fnk::scratch_workspace) and no allocation: the flagged indices go to
    // the device through a stack buffer, synchronously, a chunk at a time.
    void group_force_noalloc(const int *group, double *out) override
    {
        if (!group || !out) throw Error(FN_ERR_ARG, "group and output are required");
        if (N_ <= 0) throw Error(FN_ERR_ARG, "no bead buffer");
        const int nn = n_networks();
        if (nn > fnk::kScratchNets)
            throw Error(FN_ERR_CAPACITY, fmt("%d networks: the fixed workspace holds %d", nn, fnk::kScratchNets));
        int total = 0;
        for (int i = 0; i < N_; ++i) total += group[i] != 0;
        if (total > fnk::kScratchGroup)
            throw Error(FN_ERR_CAPACITY, fmt("%d flagged beads: the fixed workspace holds %d", total,
                                             fnk::kScratchGroup));
        fnk::ScratchWorkspace w;
        HIPCK(fnk::scratch_workspace(device, &w));
        constexpr int kChunk = 4096;
        int buf[kChunk];
        int off[fnk::kScratchNets + 1];
        int nb = 0, written = 0, max_len = 0;
        auto flush = [&] {
            if (nb == 0) return;
            HIPCK(hipMemcpyAsync(w.idx + written, buf, sizeof(int) * nb, hipMemcpyHostToDevice, stream_));
            HIPCK(hipStreamSynchronize(stream_));
            written += nb;
            nb = 0;
        };
        off[0] = 0;
        for (int k = 0; k < nn; ++k) {
            const int b0 = ens_boxes_.empty() ? 0 : ens_b0_[k], b1 = ens_boxes_.empty() ? N_ : ens_b0_[k + 1];
            for (int i = b0; i < b1; ++i)
                if (group[i] != 0) {
                    buf[nb++] = i;
                    if (nb == kChunk) flush();
                }
            off[k + 1] = written + nb;
            max_len = std::max(max_len, off[k + 1] - off[k]);
        }
        flush();
        HIPCK(hipMemcpyAsync(w.off, off, sizeof(int) * (nn + 1), hipMemcpyHostToDevice, stream_));
        fnk::group_force(frc_[fc_][0].p, w.idx, w.off, nn, max_len, w.part, w.out, stream_);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(out, w.out, sizeof(double) * nn, hipMemcpyDeviceToHost, stream_));
        HIPCK(hipStreamSynchronize(stream_));
    }

    // The rebuild of the adjoint's runs: kept candidates when they are
    // switched on (verlet 1), the one-pass exact rebuild otherwise.  Either
    // way the lists depend on the positions alone.  With kept candidates,
    // every segment's first rebuild builds them afresh, in the forward run
    // and in the segment's recomputation alike, so both build them at the
    // same steps from the same positions (and need the same capacities).
    struct AdjointRebuild {
        Engine &e;
        int saved;
        AdjointRebuild(Engine &en, bool keep) : e(en), saved(en.verlet_) { e.verlet_ = keep ? 1 : 0; }
        ~AdjointRebuild()
        {
            e.verlet_ = saved;
            e.cand_dirty_ = true;
        }
    };
    void fresh_candidates()
    {
        if (keep_cand_) fnk::fill_int(need_.p, 1, 1, stream_);
    }

    // Kept candidates at a checkpoint of an adjoint run: their CSR and the
    // positions they were built at.  Between rebuilds the build flag is 0
    // (every rebuild's filter clears it), so it is not part of the copy.
    struct CandSnap {
        DArr<int> start, list;
        DArr<double> ref[3];
        int count = 0;
    };
    void d2d(void *dst, const void *src, size_t bytes)
    {
        if (bytes) HIPCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_));
    }
    void snapshot_candidates()
    {
        int total = 0;
        HIPCK(hipMemcpyAsync(&total, cstart_.p + N_, sizeof(int), hipMemcpyDeviceToHost, stream_));
        HIPCK(hipStreamSynchronize(stream_));
        auto c = std::make_unique<CandSnap>();
        c->count = std::max(0, std::min(total, cap_));
        c->start.alloc((size_t)N_ + 1);
        c->list.alloc(std::max(c->count, 1));
        d2d(c->start.p, cstart_.p, sizeof(int) * ((size_t)N_ + 1));
        d2d(c->list.p, cand_.p, sizeof(int) * (size_t)c->count);
        for (int q = 0; q < 3; ++q) {
            c->ref[q].alloc(N_);
            d2d(c->ref[q].p, ref_[q].p, sizeof(double) * N_);
        }
        cand_ck_.push_back(std::move(c));
    }
    void restore_candidates(const CandSnap &c)
    {
        d2d(cstart_.p, c.start.p, sizeof(int) * ((size_t)N_ + 1));
        d2d(cand_.p, c.list.p, sizeof(int) * (size_t)c.count);
        for (int q = 0; q < 3; ++q) d2d(ref_[q].p, c.ref[q].p, sizeof(double) * N_);
        fnk::fill_int(need_.p, 0, 1, stream_);
    }

    void adjoint_forward(const fn_adjoint_spec &sp, const int *group, const int *steps, int ns,
                         double *samples) override
    {
        adj_.valid = false;
        const int n = sp.n_steps, guf = sp.grid_update_freq;
        if (n <= 0) throw Error(FN_ERR_ARG, "loop size should be a positive number");
        if (guf <= 0) throw Error(FN_ERR_ARG, "grid_update_freq should be positive");
        int every = sp.checkpoint_every;
        if (every < 0 || (every > 0 && every % guf != 0))
            throw Error(FN_ERR_ARG, "checkpoint_every must be 0 or a positive multiple of grid_update_freq");
        if (every == 0) every = guf * std::max(1, (int)std::lround(std::sqrt((double)n) / guf));
        if (ns < 0 || (ns > 0 && (!steps || !samples || !group)))
            throw Error(FN_ERR_ARG, "samples need a bead group, their steps and an output array");
        for (int k = 0; k < ns; ++k)
            if (steps[k] < 0 || steps[k] >= n || (k > 0 && steps[k] <= steps[k - 1]))
                throw Error(FN_ERR_ARG, "sample steps must increase strictly within [0, n_steps)");
        bool spr = sp.run_springs != 0, ang = sp.run_angles != 0;
        adj_.keep = verlet_ == 1;
        AdjointRebuild rebuild_mode(*this, adj_.keep);
        prep_fast(n, guf, spr, ang);
        std::vector<int> g(N_, 0), gi;
        if (group)
            for (int i = 0; i < N_; ++i)
                if (group[i] != 0) {
                    g[i] = 1;
                    gi.push_back(i);
                }
        group_.put(g, stream_);
        group_idx_.put(gi, stream_);
        // the group's entries of each network (networks hold consecutive beads)
        const int nn = n_networks();
        std::vector<int> goff(nn + 1, 0);
        int max_len = 0;
        for (int k = 0; k < nn; ++k) {
            const int b1 = ens_boxes_.empty() ? N_ : ens_b0_[k + 1];
            goff[k + 1] = (int)(std::lower_bound(gi.begin(), gi.end(), b1) - gi.begin());
            max_len = std::max(max_len, goff[k + 1] - goff[k]);
        }
        group_off_.put(goff, stream_);
        part_.ensure(fnk::kSumBlocks * (size_t)std::max<size_t>(std::max<size_t>(h_pairs_.size(), nn), 1));
        const int nseg = (n + every - 1) / every;
        const size_t cb = parts_bytes(state_parts(pc_, fc_, false));
        ck_.ensure(cb * nseg);
        samp_.ensure(std::max((size_t)ns * nn, (size_t)1));
        adj_.ck_pc.assign(nseg, 0);
        adj_.ck_fc.assign(nseg, 0);
        // Kept candidates carry over from one segment to the next; the
        // recomputation of a segment restores the ones its start held (a copy
        // per candidate build), so both runs build them at the same steps from
        // the same positions.  Only the first segment builds them afresh.
        adj_.ck_cand.assign(nseg, -1);
        presize_rows();
        // A run whose search rows or lists overflow is invalid: it is run
        // again from its start (the first checkpoint) with the grown
        // capacities, as run_steps does.
        for (int attempt = 0;; ++attempt) {
            cand_ck_.clear();
            int snap_builds = -1;
            if ((long long)last_total_ * 3 / 2 + 1024 > (long long)cap_) grow_list((long long)last_total_ * 2 + 1024);
            fnk::fill_int(stats_.p, 0, fnk::kStatCount, stream_);
            int si = 0;
            for (int seg = 0; seg < nseg; ++seg) {
                const int c = seg * every, len = std::min(every, n - c);
                if (seg > 0 || attempt == 0) {
                    adj_.ck_pc[seg] = pc_;
                    adj_.ck_fc[seg] = fc_;
                    save_parts(state_parts(pc_, fc_, false), ck_.p + cb * seg);
                }
                if (seg == 0 || !keep_cand_) {
                    fresh_candidates();
                } else {
                    int builds = 0;
                    HIPCK(hipMemcpyAsync(&builds, stats_.p + fnk::kStatCandBuilds, sizeof(int),
                                         hipMemcpyDeviceToHost, stream_));
                    HIPCK(hipStreamSynchronize(stream_));
                    if (builds != snap_builds) {
                        snapshot_candidates();
                        snap_builds = builds;
                    }
                    adj_.ck_cand[seg] = (int)cand_ck_.size() - 1;
                }
                const StepHook hook = [&](int k, int, int fw) {
                    if (si < ns && steps[si] == c + k) {
                        fnk::group_force(frc_[fw][0].p, group_idx_.p, group_off_.p, nn, max_len, part_.p,
                                         samp_.p + (size_t)si * nn, stream_);
                        ++si;
                    }
                };
                const bool first = seg == 0, last = seg == nseg - 1;
                enqueue_chunk(len, guf, spr, ang, first, last, nullptr, &hook);
                HIPCK(hipGetLastError());
                flip(len, first, last);
            }
            HIPCK(hipStreamSynchronize(stream_));
            try {
                check_capacity();
                break;
            } catch (const Error &e) {
                if (e.code != FN_ERR_CAPACITY || attempt == 3) throw;
            }
            pc_ = adj_.ck_pc[0];
            fc_ = adj_.ck_fc[0];
            load_parts(state_parts(pc_, fc_, false), ck_.p);
        }
        steps_ += n;
        rebuilds_ += (n + guf - 1) / guf;
        const std::vector<Part> fin = state_parts(pc_, fc_, true);
        fin_.ensure(parts_bytes(fin));
        save_parts(fin, fin_.p);
        std::vector<double> sv;
        samp_.get(sv, (size_t)ns * nn, stream_);
        HIPCK(hipStreamSynchronize(stream_));
        for (int k = 0; k < nn; ++k)          // samples[network][sample]
            for (int q = 0; q < ns; ++q) samples[(size_t)k * ns + q] = sv[(size_t)q * nn + k];
        adj_.n = n;
        adj_.guf = guf;
        adj_.every = every;
        adj_.nseg = nseg;
        adj_.spr = spr;
        adj_.ang = ang;
        adj_.n_spr = n_spr_;
        adj_.n_ang = n_ang_;
        adj_.cb = cb;
        adj_.cap = cap_;
        adj_.fin_pc = pc_;
        adj_.fin_fc = fc_;
        adj_.steps.assign(steps, steps + ns);
        adj_.n_net = nn;
        adj_.valid = true;
    }

    int n_networks() const { return ens_boxes_.empty() ? 1 : (int)ens_boxes_.size(); }

    void adjoint_backward(const double *dloss, double *gs, double *ga, double *gp, double *x0b,
                          double *w0b) override
    {
        if (!adj_.valid)
            throw Error(FN_ERR_ARG, "no fn_adjoint_forward run to differentiate (or the engine changed since)");
        const int n = adj_.n, guf = adj_.guf, every = adj_.every, nseg = adj_.nseg;
        const bool spr = adj_.spr, ang = adj_.ang;
        const int ns = (int)adj_.steps.size();
        if (ns && !dloss) throw Error(FN_ERR_ARG, "dloss is required");
        if (cap_ != adj_.cap) throw Error(FN_ERR_ARG, "the contact list was resized after the forward run");
        const int nn = adj_.n_net;
        const bool ens = !ens_boxes_.empty();
        // seeds: dL/d(sample) of network k at sample step q, as one row of nn
        // values per sampled step (dloss is [network][sample])
        std::vector<int> row_of(n, -1);
        std::vector<double> seedm((size_t)std::max(ns, 1) * nn, 0.0);
        for (int q = 0; q < ns; ++q) {
            row_of[adj_.steps[q]] = q;
            for (int k = 0; k < nn; ++k) seedm[(size_t)q * nn + k] = dloss[(size_t)k * ns + q];
        }
        seed_.put(seedm, stream_);
        auto seed_at = [&](int step) -> const double * {
            return (step >= 0 && row_of[step] >= 0) ? seed_.p + (size_t)row_of[step] * nn : nullptr;
        };
        AdjointRebuild rebuild_mode(*this, adj_.keep);
        if (ens) {
            prepare_ensemble_run(spr, ang);
        } else {
            n_spr_ = adj_.n_spr;
            n_ang_ = adj_.n_ang;
            geometry(settings_of(K_FILL_GRID), settings_of(K_BUILD_LIST));
        }
        const size_t N = N_;
        const int np = (int)h_pairs_.size();
        // contact-scale columns: the pair entries (single network), or the
        // largest pair table of a network (ensemble, entries local to it)
        int cols = np;
        if (ens) {
            cols = 0;
            for (int k = 0; k < nn; ++k) {
                const int e1 = k + 1 < nn ? ens_boxes_[k + 1].pair_base : np;
                cols = std::max(cols, e1 - ens_boxes_[k].pair_base);
            }
        }
        for (auto *d : {&fb_[0], &fb_[1], &xb_, &wb_}) d->ensure(N);
        gsp_.ensure(std::max(S_, 1));
        gan_.ensure(std::max(A_, 1));
        const bool chain_adj = chain_ok_ && chain_on_ && !ens;
        const int prow = fnk::adjoint_pair_rows(N_, cols, ens, chain_adj);
        gpr_.ensure(std::max<size_t>((size_t)prow * cols, 1));
        gps_.ensure(std::max(np, 1));
        if (!fnk::adjoint_one_launch(chain_adj, ens, cols)) {
            sp_jv_.ensure(std::max<size_t>(2 * (size_t)S_, 1));
            an_jv_.ensure(std::max<size_t>(3 * (size_t)A_, 1));
        }
        HIPCK(hipMemsetAsync(xb_.p, 0, sizeof(double3) * N, stream_));
        HIPCK(hipMemsetAsync(wb_.p, 0, sizeof(double3) * N, stream_));
        HIPCK(hipMemsetAsync(gsp_.p, 0, sizeof(double) * gsp_.n, stream_));
        HIPCK(hipMemsetAsync(gan_.p, 0, sizeof(double) * gan_.n, stream_));
        HIPCK(hipMemsetAsync(gpr_.p, 0, sizeof(double) * gpr_.n, stream_));
        // per-step records of one segment: positions (and images), bond
        // statuses; per rebuild: the contact list
        const int nimg = alias_ ? 3 : 6;
        const size_t SA = (size_t)S_ + A_;
        const int nr = (every + guf - 1) / guf;
        const size_t lrow = 2 * N + 1 + (size_t)cap_;
        rec_pos_.ensure((size_t)every * nimg * N);
        rec_st_.ensure(std::max<size_t>((size_t)every * SA, 1));
        rec_nb_.ensure((size_t)nr * lrow);
        fnk::adjoint_seed<T>(fb_[0].p, wb_.p, hlf_.p, group_.p, seed_at(n - 1), ens_net(), N_, stream_);
        int cur = 0;
        for (int seg = nseg - 1; seg >= 0; --seg) {
            const int c = seg * every, len = std::min(every, n - c);
            pc_ = adj_.ck_pc[seg];
            fc_ = adj_.ck_fc[seg];
            load_parts(state_parts(pc_, fc_, false), ck_.p + adj_.cb * seg);
            if (adj_.ck_cand[seg] < 0)
                fresh_candidates();
            else
                restore_candidates(*cand_ck_[adj_.ck_cand[seg]]);
            const StepHook hook = [&](int k, int pr, int) {
                const double *p[3] = {pos_[pr][0].p, pos_[pr][1].p, pos_[pr][2].p};
                const double *im[3] = {img_[pr][0].p, img_[pr][1].p, img_[pr][2].p};
                fnk::record_step(p, alias_ ? nullptr : im, sp_status_.p, S_, an_status_.p, A_, N_,
                                 rec_pos_.p + (size_t)k * nimg * N, rec_st_.p + (size_t)k * SA, stream_);
                if (k % guf == 0) {
                    int *dl = rec_nb_.p + (size_t)(k / guf) * lrow;
                    d2d(dl, nb_start_.p, sizeof(int) * (N + 1));
                    d2d(dl + N + 1, nb_cnt_.p, sizeof(int) * N);
                    fnk::copy_bounded(nb_list_.p, dl + 2 * N + 1, nb_start_.p + N, cap_, stream_);
                }
            };
            enqueue_chunk(len, guf, spr, ang, seg == 0, seg == nseg - 1, nullptr, &hook);
            for (int k = len - 1; k >= 0; --k) {
                const P *rp = rec_pos_.p + (size_t)k * nimg * N;
                const P *ri = alias_ ? rp : rp + 3 * N;
                const unsigned char *rs = rec_st_.p + (size_t)k * SA;
                const int *rl = rec_nb_.p + (size_t)(k / guf) * lrow;
                fnk::AdjArgs<T, P> a{};
                a.n = N_;
                a.n_springs = spr ? n_spr_ : 0;
                a.n_angles = ang ? n_ang_ : 0;
                a.springs = spr ? S_ : 0;
                a.angles = ang ? A_ : 0;
                a.sp_jv = sp_jv_.p;
                a.an_jv = an_jv_.p;
                a.drift = kbox(K_ADVANCE_POSITION);
                a.force = kbox(K_RESOLVE_CONTACTS);
                a.net = ens_net();
                a.boxes = boxes_.p;
                a.px = rp; a.py = rp + N; a.pz = rp + 2 * N;
                a.ix = ri; a.iy = ri + N; a.iz = ri + 2 * N;
                a.lam = lam_.p; a.hlf = hlf_.p; a.hlf2 = hlf2_.p; a.radius = rad_.p;
                a.sp_start = sp_start_.p; a.sp_slot = sp_slot_.p; a.sp_p1 = sp_p1_.p; a.sp_p2 = sp_p2_.p;
                a.sp_type = sp_type_.p; a.sp_on = rs; a.sp_l0 = sp_l0_.p; a.sp_k = sp_k_.p;
                a.an_start = an_start_.p; a.an_slot = an_slot_.p; a.an_b1 = an_b1_.p; a.an_b2 = an_b2_.p;
                a.an_b3 = an_b3_.p; a.an_type = an_type_.p; a.an_on = rs + S_; a.an_th0 = an_th0_.p; a.an_k = an_k_.p;
                a.nb_start = rl; a.nb_cnt = rl + N + 1; a.nb_list = rl + 2 * N + 1;
                a.pairs = pairs_.p;
                a.n_pairs = cols;
                a.xb = xb_.p; a.wb = wb_.p;
                a.fb = fb_[cur].p; a.fb_next = fb_[1 - cur].p;
                a.group = group_.p;
                a.seed_prev = seed_at(c + k - 1);
                a.g_spring = gsp_.p; a.g_angle = gan_.p; a.g_pair = gpr_.p;
                if (chain_adj) {
                    a.ch_spring = ch_spring_.p;
                    a.ch_angle = ch_angle_.p;
                    if (cmp_ok_) {
                        a.bond_code = bond_code_.p;
                        a.sp_sets = sp_sets_.p;
                        a.an_sets = an_sets_.p;
                    }
                }
                fnk::adjoint_step(a, stream_);
                cur ^= 1;
            }
            HIPCK(hipGetLastError());
        }
        // back to the state the forward run ended in
        pc_ = adj_.fin_pc;
        fc_ = adj_.fin_fc;
        load_parts(state_parts(pc_, fc_, true), fin_.p);
        if (ens) {
            std::vector<int> boff(ens_b0_.begin(), ens_b0_.end()), pbase(nn + 1, np);
            for (int k = 0; k < nn; ++k) pbase[k] = ens_boxes_[k].pair_base;
            net_off_.put(boff, stream_);
            pair_off_.put(pbase, stream_);
            fnk::sum_network_columns(gpr_.p, cols, net_off_.p, pair_off_.p, nn, gps_.p, stream_);
        } else {
            fnk::sum_columns(gpr_.p, prow, np, part_.p, gps_.p, stream_);
        }
        auto d2h = [&](void *dst, const void *src, size_t bytes) {
            if (dst && bytes) HIPCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_));
        };
        d2h(gs, gsp_.p, sizeof(double) * S_);
        d2h(ga, gan_.p, sizeof(double) * A_);
        d2h(gp, gps_.p, sizeof(double) * np);
        d2h(x0b, xb_.p, sizeof(double3) * N);
        d2h(w0b, wb_.p, sizeof(double3) * N);
        HIPCK(hipStreamSynchronize(stream_));
    }

    struct AdjRun {
        bool valid = false, spr = false, ang = false, keep = false;
        int n = 0, guf = 1, every = 1, nseg = 0, n_spr = 0, n_ang = 0, cap = 0, fin_pc = 0, fin_fc = 0, n_net = 1;
        size_t cb = 0;
        std::vector<int> ck_pc, ck_fc, ck_cand, steps;   // ck_cand: kept candidates' copy (-1: built afresh)
    };

    void run_chunk(int len, int guf, bool spr, bool ang, bool first, bool last)
    {
        const auto key = std::make_tuple(len, guf, (int)spr, (int)ang, (int)first, (int)last, pc_, fc_, n_spr_, n_ang_);
        if (graphs_on_) {
            auto it = graphs_.find(key);
            if (it == graphs_.end()) {
                GraphRec rec;
                hipGraph_t g;
                HIPCK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                capturing_ = true;
                enqueue_chunk(len, guf, spr, ang, first, last, timing_ ? &rec.ev : nullptr);
                capturing_ = false;
                HIPCK(hipStreamEndCapture(stream_, &g));
                HIPCK(hipGraphInstantiate(&rec.ex[0], g, nullptr, nullptr, 0));
                HIPCK(hipGraphInstantiate(&rec.ex[1], g, nullptr, nullptr, 0));
                // both copies go to the device now, so that neither one's
                // first replay waits for its upload (where the runtime cannot,
                // the first replay uploads it: no error is kept)
                for (hipGraphExec_t x : rec.ex)
                    if (hipGraphUpload(x, stream_) != hipSuccess) (void)hipGetLastError();
                HIPCK(hipGraphDestroy(g));
                it = graphs_.emplace(key, rec).first;
            }
            GraphRec &gr = it->second;
            HIPCK(hipGraphLaunch(gr.ex[gr.next], stream_));
            gr.next ^= 1;
            if (timing_) {   // the graph's event nodes hold this replay's times until the next replay
                HIPCK(hipStreamSynchronize(stream_));
                accumulate(it->second.ev);
            }
        } else {
            loose_.emplace_back();
            enqueue_chunk(len, guf, spr, ang, first, last, timing_ ? &loose_.back() : nullptr);
            if (timing_) used_.push_back(&loose_.back());
        }
        HIPCK(hipGetLastError());
        flip(len, first, last);
    }

    // Buffer parity after a chunk: the first drift and every step but the
    // run's last flip the positions, every step flips the forces.
    void flip(int len, bool first, bool last)
    {
        const int flips_p = (first ? 1 : 0) + len - (last ? 1 : 0);
        pc_ ^= (flips_p & 1);
        fc_ ^= (len & 1);
        grid_valid_ = true;
    }

    void clear_graphs()
    {
        for (auto &kv : graphs_) {
            for (hipGraphExec_t x : kv.second.ex)
                if (x) (void)hipGraphExecDestroy(x);
            kv.second.ev.release();
        }
        graphs_.clear();
        for (auto &e : loose_) e.release();
        loose_.clear();
        used_.clear();
    }

    // After a synchronised run: fold the events of every chunk into the totals
    // (a graph replayed several times holds the timings of its last replay).
    void collect_timing()
    {
        if (!timing_) return;
        std::vector<const Events *> seen;
        for (const Events *e : used_)
            if (std::find(seen.begin(), seen.end(), e) == seen.end()) {
                seen.push_back(e);
                accumulate(*e);
            }
        used_.clear();
        for (auto &e : loose_) e.release();
        loose_.clear();
    }

    // A captured chunk, instantiated twice: consecutive replays of one chunk
    // alternate between the two, so that a replay is submitted while the
    // previous one still runs (the runtime holds back a second launch of an
    // executable until its first has finished).
    struct GraphRec {
        hipGraphExec_t ex[2] = {nullptr, nullptr};
        int next = 0;
        Events ev;
    };

    hipStream_t stream_ = nullptr;
    std::vector<Buf> bufs_;
    int role_[10];
    std::map<std::string, KernelRec> kernels_;
    std::map<std::tuple<int, int, int, int, int, int, int, int, int, int>, GraphRec> graphs_;
    std::list<Events> loose_;
    std::vector<const Events *> used_;
    bool timing_ = false;
    bool capturing_ = false;
    fn_timing timing_acc_{};

    int N_ = 0, S_ = 0, A_ = 0, C_ = 0;
    int pc_ = 0, fc_ = 0;
    int n_spr_ = 0, n_ang_ = 0;
    bool alias_ = false, pos_eq_img_ = false, csr_dirty_ = true, grid_valid_ = false, nb_init_ = false;
    bool cellnh_checked_ = false;
    int cap_ = 0, last_total_ = 0;
    int K_ = 64;   // search-row length: entries one bead may collect before compaction
    int rows_exact_ = 0, rows_cand_ = 0;   // longest exact / candidate rows seen since the state was set (0: none)
    DArr<unsigned char> run_snap_;         // the state at the start of a guarded run
    int chunk_ = 128;                                // steps per captured graph
    int block_ = 256;                                // threads per block of the bead kernels
    bool chain_on_ = true;                           // use the chain bead kernel where the layout allows
    bool chain_ok_ = false;                          // the bonds follow the chain layout
    int verlet_ = 2;                                 // candidates kept between rebuilds: 0 never, 1 always, 2 auto
    bool verlet_trial_ = false;                      // auto: candidates currently kept
    int verlet_runs_ = 0;                            // auto: runs since the last trial
    bool staged_ = false;                            // neighbour search reads its block's neighbourhood from LDS
    bool flat_ = true;                               // 3x3x3 neighbour search flattened over each wave
    int skin_permille_ = 300;                        // skin as a fraction of the listing reach
    bool keep_cand_ = false, cand_dirty_ = true;
    int *h_stats_ = nullptr;                         // pinned copy of the statistics at the end of a run
    double cand_r2_ = -1.0, lim2_ = 0.0;
    fnk::Stencil stencil_{}, stencil_exact_{};
    bool graphs_on_ = true;
    long long steps_ = 0, rebuilds_ = 0;

    DArr<P> pos_[2][3], img_[2][3];
    DArr<T> vel_[3], vold_[3], frc_[2][3];
    DArr<T> lam_, hlf_, hlf2_, rad_;
    DArr<int> type_, tag_, loc_;
    DArr<double> rad_d_, skin_d_;
    DArr<int> sp_p1_, sp_p2_, sp_type_, sp_status_, sp_break_, sp_start_, sp_slot_;
    DArr<double> sp_l0_, sp_c1_, sp_c2_;
    DArr<T> sp_k_;
    DArr<int> an_b1_, an_b2_, an_b3_, an_s1_, an_s2_, an_type_, an_status_, an_start_, an_slot_;
    DArr<T> an_th0_, an_k_;
    DArr<fnk::V3<T>> sp_term_, an_term_;   // bond terms of the current step at the incidence slots
    DArr<int> ch_spring_, ch_angle_;       // chain layout: spring (i, i+1), angle centred on i
    bool cmp_ok_ = false;                  // the chain layout's bond parameters by bead are valid
    bool ms_ok_ = false;                   // the beads' mass sets are valid
    int n_ms_ = 0;
    DArr<unsigned char> ms_code_;          // per bead: its mass set
    DArr<fnk::MassSet<T>> ms_sets_;
    DArr<unsigned> bond_code_;             // per bead: the sets and statuses of spring (i, i+1) and angle at i
    DArr<fnk::SpringSet> sp_sets_;
    DArr<fnk::AngleSet> an_sets_;
    AdjRun adj_;                           // the last adjoint forward run
    DArr<unsigned char> ck_, fin_;         // its checkpoints (one per segment) and end state
    std::vector<std::unique_ptr<CandSnap>> cand_ck_;   // and the kept candidates its segments start from
    DArr<int> group_, group_idx_;          // sampled beads: flags, ascending indices
    DArr<int> group_off_, net_off_, pair_off_;   // per network: group entries, beads, pair entries
    DArr<double> seed_;                    // dL/d(sample) rows of the sampled steps
    DArr<double> part_;                    // partial sums of the deterministic reductions
    DArr<double3> sp_jv_, an_jv_;          // bond terms of an adjoint step at the incidence slots
    DArr<double> samp_;                    // samples of the run
    DArr<double3> fb_[2], xb_, wb_;        // adjoint state
    DArr<double> gsp_, gan_, gpr_, gps_;   // scale adjoints (contacts per bead and pair entry, then summed)
    DArr<P> rec_pos_;                      // per step of a segment: positions and images of the force evaluation
    DArr<unsigned char> rec_st_;           // per step: the bonds that act
    DArr<int> rec_nb_;                     // per rebuild: the contact list
    DArr<int> rf_idx_, rf_off_;            // fn_group_force: flagged beads, their entries per network
    DArr<double> rf_part_, rf_out_;        // its partial sums and result
    DArr<fnk::PairC> pairs_;
    DArr<int> cstart_;   // offsets of the kept candidates' rows (nb_start_: of the contact list)
    DArr<int> key_, ckey_, rank_, cell_count_, cell_start_, ncnt_, nb_start_, nb_cnt_, cand_, nb_list_, scan_tmp_,
        stats_, step_, need_;
    DArr<double> ref_[3];   // positions at the last candidate build
    // ensemble: Box and settings of every network, first bead / spring / angle
    // of every network (and the totals), network of every bead
    double max_radius_ = 0.0, max_skin_ = 0.0;   // over all beads: the listing reach of hs pairs
    int types_checked_ = -1;                     // types_num the bead types were last checked against
    std::vector<fnk::Box> ens_boxes_;
    std::vector<fn_settings> ens_sets_;
    std::vector<int> ens_b0_, ens_s0_, ens_a0_;
    bool ens_finite_ = false;
    bool ens_grids3_ = false;    // every network's grid has at least three cells per axis
    DArr<fnk::Box> boxes_;
    DArr<int> net_;
    DArr<fnk::SortedBead> so_bead_;
    DArr<fnk::CellBead> cellbead_;   // the filter's rounded positions and cells (kept candidates)
    fnk::CellPack pack_{10, 20};       // bit offsets of the cell coordinates in cellbead_
    DArr<int> scratch_;
    double reach2_ = 0.0;

    std::vector<fn_bead_info> h_info_;
    std::vector<fn_spring> h_springs_;
    std::vector<fn_angle> h_angles_;
    std::vector<fn_pair> h_pairs_;
    std::vector<fn_grid_cell_nh> h_cellnh_;
};

template <typename F>
int guarded(fn_engine *eng, F &&f)
{
    if (!eng) return FN_ERR_ARG;
    try {
        HIPCK(hipSetDevice(eng->device));
        f();
        return FN_OK;
    } catch (const Error &e) {
        eng->err = e.what();
        eng->log(e.what());
        return e.code;
    } catch (const std::exception &e) {
        eng->err = e.what();
        eng->log(e.what());
        return FN_ERR_ARG;
    }
}

}  // namespace

// ---- C ABI ---------------------------------------------------------------------------------------
extern "C" {

const char *fn_version(void) { return "fibrenet_amd 0.1 (gfx950)"; }

int fn_create(int device, int precision, fn_engine **out)
{
    if (!out) return FN_ERR_ARG;
    *out = nullptr;
    try {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return FN_ERR_HIP;
        if (device < 0 || device >= n) return FN_ERR_ARG;
        if (precision == FN_PREC_F64) *out = new Engine<double, double>(device);
        else if (precision == FN_PREC_F32) *out = new Engine<float, double>(device);
        else return FN_ERR_ARG;
        return FN_OK;
    } catch (const Error &e) {
        fprintf(stderr, "fibrenet: %s\n", e.what());
        return e.code;
    } catch (...) {
        return FN_ERR_HIP;
    }
}

int fn_destroy(fn_engine *eng)
{
    if (!eng) return FN_ERR_ARG;
    (void)hipSetDevice(eng->device);
    delete eng;
    return FN_OK;
}

const char *fn_last_error(fn_engine *eng) { return eng ? eng->err.c_str() : "null engine"; }

int fn_set_log_callback(fn_engine *eng, fn_log_fn fn, void *user)
{
    if (!eng) return FN_ERR_ARG;
    eng->log_fn = fn;
    eng->log_user = user;
    return FN_OK;
}

int fn_register(fn_engine *eng, const char *const *names, int n)
{
    return guarded(eng, [&] {
        if (!names || n <= 0) throw Error(FN_ERR_ARG, "no kernel names");
        for (int i = 0; i < n; ++i) eng->reg(names[i]);
    });
}

int fn_buffer_create(fn_engine *eng, int kind, int elem_size, int n_elems, const void *host, int *id)
{
    return guarded(eng, [&] {
        // the pair table of one network (an ensemble checks each network's own)
        if (kind == FN_BUF_PAIR && n_elems > (1 << (31 - fnk::kPidShift)))
            throw Error(FN_ERR_UNSUPPORTED, fmt("a pair table of %d entries: the engine holds at most %d "
                                                "(at most 7 bead types)", n_elems, 1 << (31 - fnk::kPidShift)));
        const int r = eng->create_buffer(kind, elem_size, n_elems, host);
        if (id) *id = r;
    });
}

int fn_buffer_write(fn_engine *eng, int id, const void *host, int n_padded, int elem_size)
{
    return guarded(eng, [&] { eng->write_buffer(id, host, n_padded, elem_size); });
}

int fn_buffer_read(fn_engine *eng, int id, void *host)
{
    return guarded(eng, [&] {
        if (!host) throw Error(FN_ERR_ARG, "null destination");
        eng->read_buffer(id, host);
    });
}

int fn_buffer_info(fn_engine *eng, int id, int *size_bytes, int *elements, int *length, int *struct_size)
{
    return guarded(eng, [&] {
        const Buf *b = eng->buffer(id);
        if (!b) throw Error(FN_ERR_ARG, fmt("no buffer %d", id));
        if (size_bytes) *size_bytes = b->elements * b->esize;
        if (elements) *elements = b->elements;
        if (length) *length = b->length;
        if (struct_size) *struct_size = b->esize;
    });
}

int fn_set_arg(fn_engine *eng, const char *kernel, int index, int arg_kind, const void *value, int size)
{
    return guarded(eng, [&] { eng->set_arg(kernel, index, arg_kind, value, size); });
}

int fn_set_range(fn_engine *eng, const char *kernel, int range)
{
    return guarded(eng, [&] { eng->set_range(kernel, range); });
}

int fn_launch(fn_engine *eng, const char *kernel)
{
    return guarded(eng, [&] { eng->launch(kernel); });
}

int fn_finish(fn_engine *eng)
{
    return guarded(eng, [&] { eng->finish(); });
}

int fn_run_steps(fn_engine *eng, int n_steps, int grid_update_freq, int run_springs, int run_angles)
{
    return guarded(eng, [&] { eng->run_steps(n_steps, grid_update_freq, run_springs != 0, run_angles != 0); });
}

int fn_export_cells(fn_engine *eng, int *cell_of_bead)
{
    return guarded(eng, [&] {
        if (!cell_of_bead) throw Error(FN_ERR_ARG, "null destination");
        eng->export_cells(cell_of_bead);
    });
}

int fn_export_pairs(fn_engine *eng, int *pairs, long long capacity, long long *n_pairs)
{
    return guarded(eng, [&] {
        const long long n = eng->export_pairs(pairs, capacity);
        if (n_pairs) *n_pairs = n;
    });
}

int fn_export_breaks(fn_engine *eng, int *break_step)
{
    return guarded(eng, [&] {
        if (!break_step) throw Error(FN_ERR_ARG, "null destination");
        eng->export_breaks(break_step);
    });
}

int fn_set_option(fn_engine *eng, const char *name, int value)
{
    return guarded(eng, [&] { eng->set_option(name ? name : "", value); });
}

int fn_get_option(fn_engine *eng, const char *name, int *value)
{
    return guarded(eng, [&] {
        const int v = eng->get_option(name ? name : "");
        if (value) *value = v;
    });
}

int fn_set_timing(fn_engine *eng, int on)
{
    return guarded(eng, [&] { eng->set_timing(on != 0); });
}

int fn_copy_probe(fn_engine *eng, long long bytes, double *ms)
{
    return guarded(eng, [&] {
        if (bytes <= 0 || !ms) throw Error(FN_ERR_ARG, "copy probe needs a positive size and a destination");
        eng->copy_probe(bytes, ms);
    });
}

int fn_batch_create(int device, int precision, int n_networks, const fn_network *nets, fn_engine **out)
{
    const int rc = fn_create(device, precision, out);
    if (rc != FN_OK) return rc;
    const int lr = guarded(*out, [&] { (*out)->load_ensemble(nets, n_networks); });
    return lr;
}

int fn_batch_load(fn_engine *eng, int n_networks, const fn_network *nets)
{
    return guarded(eng, [&] { eng->reload_ensemble(nets, n_networks); });
}

int fn_batch_read(fn_engine *eng, int network, int kind, void *out)
{
    return guarded(eng, [&] {
        if (!out) throw Error(FN_ERR_ARG, "null destination");
        eng->ensemble_read(network, kind, out);
    });
}

int fn_autotune(fn_engine *eng, int n_steps, int grid_update_freq, int run_springs, int run_angles,
                fn_tune_trial *trials, int capacity, int *n_trials)
{
    return guarded(eng, [&] {
        std::vector<fn_tune_trial> t;
        eng->autotune(n_steps, grid_update_freq, run_springs != 0, run_angles != 0, t);
        if (n_trials) *n_trials = (int)t.size();
        if (trials)
            for (int k = 0; k < (int)t.size() && k < capacity; ++k) trials[k] = t[k];
    });
}

int fn_adjoint_forward(fn_engine *eng, const fn_adjoint_spec *spec, const int *group, const int *sample_steps,
                       int n_samples, double *samples)
{
    return guarded(eng, [&] {
        if (!spec) throw Error(FN_ERR_ARG, "spec is required");
        eng->adjoint_forward(*spec, group, sample_steps, n_samples, samples);
    });
}

int fn_adjoint_backward(fn_engine *eng, const double *dloss, double *d_spring, double *d_angle, double *d_pair,
                        double *d_pos0, double *d_vel0)
{
    return guarded(eng, [&] { eng->adjoint_backward(dloss, d_spring, d_angle, d_pair, d_pos0, d_vel0); });
}

int fn_checkpoint_save(fn_engine *eng, const char *path)
{
    return guarded(eng, [&] { eng->checkpoint_save(path); });
}

int fn_checkpoint_load(fn_engine *eng, const char *path)
{
    return guarded(eng, [&] { eng->checkpoint_load(path); });
}

int fn_group_force(fn_engine *eng, const int *group, double *out)
{
    return guarded(eng, [&] { eng->group_force(group, out); });
}

int fn_group_force_noalloc(fn_engine *eng, const int *group, double *out)
{
    return guarded(eng, [&] { eng->group_force_noalloc(group, out); });
}

int fn_get_timing(fn_engine *eng, fn_timing *out)
{
    return guarded(eng, [&] {
        if (!out) throw Error(FN_ERR_ARG, "null destination");
        eng->timing(out);
    });
}

int fn_get_stats(fn_engine *eng, fn_stats *out)
{
    return guarded(eng, [&] {
        if (!out) throw Error(FN_ERR_ARG, "null destination");
        eng->stats(out);
    });
}

}  // extern "C"
