// fibrenet_manager.hpp - drop-in C++ facade with the call surface of
// FibreNet's OpenclManager (OpenclManager.h:20-71), implemented over the C ABI
// of fibrenet.h.  SimulationManager's calls keep their names, argument order
// and return conventions:
//
//   OpenclManager                     fibrenet::HipManager
//   Initialize(log)                   Initialize(log)           (device 0, or FIBRENET_DEVICE)
//   RegisterKernel(names, code)       RegisterKernel(names, code) (code is ignored: kernels are prebuilt)
//   CreateBuffer<T>(vector&, flags)   CreateBuffer<T>(vector&, flags)  pads `data` in place to 64
//   CreateBuffer<T>(int, flags)       CreateBuffer<T>(int, flags)
//   WriteBuffer<T>(id, vector&)       WriteBuffer<T>(id, vector&)       1 / 0 / -1 / -2
//   ReadBuffer(id, void*)             ReadBuffer(id, void*)
//   GetBuffer(id) -> cl::Buffer*      GetBuffer(id) -> DeviceBuffer*
//   SetKernelArg<T>(name, idx, v)     SetKernelArg<T>(name, idx, v)  (DeviceBuffer, int, settings)
//   SetKernelGlobalWorkRange          SetKernelGlobalWorkRange
//   ExecuteKernel(name | id)          ExecuteKernel(name | id)
//   RunKernels()                      RunKernels()
//   getters, GetKernelId, Close       same
//
// Record types map to buffer kinds through fibrenet::buffer_kind<T>; it is
// specialised here for the fn_* wire structs and for int.  A caller with its
// own identically laid-out structs (FibreNet's Bead, BeadInfo, ...) adds a
// one-line specialisation per struct.
#pragma once
#include <cmath>
#include <functional>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "fibrenet.h"

#ifndef CL_MEM_READ_WRITE
#define CL_MEM_READ_WRITE (1 << 0)
#define CL_MEM_WRITE_ONLY (1 << 1)
#define CL_MEM_READ_ONLY (1 << 2)
#endif

namespace fibrenet {

typedef unsigned long long mem_flags;

template <class T> struct buffer_kind { static constexpr int value = 0; };
template <> struct buffer_kind<fn_bead> { static constexpr int value = FN_BUF_BEAD; };
template <> struct buffer_kind<fn_bead_info> { static constexpr int value = FN_BUF_BEAD_INFO; };
template <> struct buffer_kind<fn_spring> { static constexpr int value = FN_BUF_SPRING; };
template <> struct buffer_kind<fn_angle> { static constexpr int value = FN_BUF_ANGLE; };
template <> struct buffer_kind<fn_pair> { static constexpr int value = FN_BUF_PAIR; };
template <> struct buffer_kind<int> { static constexpr int value = FN_BUF_INT; };
template <> struct buffer_kind<fn_grid_cell> { static constexpr int value = FN_BUF_GRID_CELL; };
template <> struct buffer_kind<fn_grid_cell_nh> { static constexpr int value = FN_BUF_GRID_CELL_NH; };
template <> struct buffer_kind<fn_particle_nh> { static constexpr int value = FN_BUF_PARTICLE_NH; };

// Handle returned by GetBuffer, passed back to SetKernelArg (cl::Buffer's role).
struct DeviceBuffer { int id; };

// Argument kinds accepted by SetKernelArg: buffers, ints and the settings
// record (any type laid out as fn_settings, e.g. SimulationSettings).
template <class T> struct arg_kind { static constexpr int value = sizeof(T) == sizeof(fn_settings) ? FN_ARG_SETTINGS : 0; };
template <> struct arg_kind<DeviceBuffer> { static constexpr int value = FN_ARG_BUFFER; };
template <> struct arg_kind<int> { static constexpr int value = FN_ARG_INT; };

class HipManager {
  public:
    typedef std::function<void(const std::string &)> log_fn;

    HipManager() = default;
    HipManager(const HipManager &) = delete;
    HipManager &operator=(const HipManager &) = delete;
    ~HipManager() { if (eng_) fn_destroy(eng_); }

    bool Initialize(log_fn log = nullptr, int precision = FN_PREC_F32, int device = -1)
    {
        log_ = log;
        if (device < 0) {
            const char *e = std::getenv("FIBRENET_DEVICE");
            device = e ? std::atoi(e) : 0;
        }
        if (fn_create(device, precision, &eng_) != FN_OK) {
            say("HipManager::Initialize: ERROR - unable to open the HIP device");
            return false;
        }
        if (log_) fn_set_log_callback(eng_, &HipManager::forward, this);
        return true;
    }

    bool GetStatus() const { return eng_ != nullptr; }
    fn_engine *engine() { return eng_; }

    bool RegisterKernel(const std::vector<std::string> &names, const std::string & /*code*/ = std::string())
    {
        if (!eng_) return false;
        std::vector<const char *> p;
        for (const auto &n : names) p.push_back(n.c_str());
        if (fn_register(eng_, p.data(), (int)p.size()) != FN_OK) return fail("RegisterKernel");
        for (const auto &n : names) {
            names_.push_back(n);
            ranges_.push_back(0);
        }
        return true;
    }

    template <class T> int CreateBuffer(std::vector<T> &data, mem_flags /*flags*/)
    {
        static_assert(buffer_kind<T>::value != 0, "specialise fibrenet::buffer_kind for this record type");
        if (data.empty()) { say("HipManager::CreateBuffer: ERROR - buffer is empty"); return -2; }
        const int n = (int)data.size();
        data.resize(pad(n));
        return create(buffer_kind<T>::value, (int)sizeof(T), n, data.data());
    }

    template <class T> int CreateBuffer(int size, mem_flags /*flags*/)
    {
        static_assert(buffer_kind<T>::value != 0, "specialise fibrenet::buffer_kind for this record type");
        if (size <= 0) { say("HipManager::CreateBuffer: ERROR - buffer is empty"); return -1; }
        return create(buffer_kind<T>::value, (int)sizeof(T), size, nullptr);
    }

    template <class T> int WriteBuffer(int id, std::vector<T> &data)
    {
        if (data.empty()) { say("HipManager::WriteBuffer: ERROR - source buffer is empty"); return -2; }
        const int rc = fn_buffer_write(eng_, id, data.data(), (int)data.size(), (int)sizeof(T));
        if (rc == FN_OK) return 1;
        if (rc == FN_ERR_MISMATCH) return 0;
        fail("WriteBuffer");
        return -1;
    }

    bool ReadBuffer(int id, void *ptr)
    {
        if (fn_buffer_read(eng_, id, ptr) != FN_OK) return fail("ReadBuffer");
        return true;
    }

    DeviceBuffer *GetBuffer(const int &id)
    {
        if (id < 0 || id >= (int)handles_.size()) return nullptr;
        return handles_[id].get();
    }

    int GetBufferSize(const int &id) { int v; return info(id, &v, nullptr, nullptr, nullptr) ? v : -1; }
    int GetBufferElements(const int &id) { int v; return info(id, nullptr, &v, nullptr, nullptr) ? v : -1; }
    int GetBufferLength(const int &id) { int v; return info(id, nullptr, nullptr, &v, nullptr) ? v : -1; }
    int GetBufferStructureSize(const int &id) { int v; return info(id, nullptr, nullptr, nullptr, &v) ? v : -1; }
    int GetLocalWorkGroupSize() const { return 64; }
    int GetNumberOfKernels() const { return names_.empty() ? -1 : (int)names_.size(); }

    template <class T> bool SetKernelArg(const std::string &name, unsigned idx, const T &v)
    {
        static_assert(arg_kind<T>::value != 0, "kernel arguments are DeviceBuffer, int or the settings record");
        int rc;
        if (arg_kind<T>::value == FN_ARG_BUFFER)
            rc = fn_set_arg(eng_, name.c_str(), (int)idx, FN_ARG_BUFFER, &reinterpret_cast<const DeviceBuffer &>(v).id,
                            (int)sizeof(int));
        else
            rc = fn_set_arg(eng_, name.c_str(), (int)idx, arg_kind<T>::value, &v, (int)sizeof(T));
        if (rc != FN_OK) return fail("SetKernelArg");
        return true;
    }

    void SetKernelGlobalWorkRange(const std::string &name, int range)
    {
        if (fn_set_range(eng_, name.c_str(), range) != FN_OK) { fail("SetKernelGlobalWorkRange"); return; }
        const int k = GetKernelId(name);
        if (k >= 0) ranges_[k] = range;
    }

    int GetKernelId(const std::string &name) const
    {
        for (int i = 0; i < (int)names_.size(); ++i)
            if (names_[i] == name) return i;
        return -1;
    }

    bool ExecuteKernel(const std::string &name)
    {
        if (fn_launch(eng_, name.c_str()) != FN_OK) return fail("ExecuteKernel");
        return true;
    }

    bool ExecuteKernel(int id)
    {
        if (id < 0 || id >= (int)names_.size()) { say("HipManager::ExecuteKernel: ERROR - kernel structure is empty"); return false; }
        return ExecuteKernel(names_[id]);
    }

    bool RunKernels()
    {
        if (fn_finish(eng_) != FN_OK) return fail("RunKernels");
        return true;
    }

    // Fast path: n steps of RunSimulationLoop in fused, graph-captured launches.
    bool RunSteps(int n, int grid_update_freq, bool springs, bool angles)
    {
        if (fn_run_steps(eng_, n, grid_update_freq, springs, angles) != FN_OK) return fail("RunSteps");
        return true;
    }

    void Close() { say("\tHipManager: manager cleanup...."); }

  private:
    static int pad(int n) { return 64 * (int)std::ceil((double)n / 64.0); }

    int create(int kind, int esize, int n, const void *host)
    {
        int id = -1;
        if (fn_buffer_create(eng_, kind, esize, n, host, &id) != FN_OK) { fail("CreateBuffer"); return -1; }
        if ((int)handles_.size() <= id) handles_.resize(id + 1);
        handles_[id].reset(new DeviceBuffer{id});
        return id;
    }

    bool info(int id, int *size, int *el, int *len, int *ss)
    {
        return eng_ && fn_buffer_info(eng_, id, size, el, len, ss) == FN_OK;
    }

    bool fail(const char *what)
    {
        say(std::string("\tHipManager::") + what + ": ERROR - " + (eng_ ? fn_last_error(eng_) : "not initialized"));
        return false;
    }

    void say(const std::string &m) { if (log_) log_(m); }

    static void forward(const char *m, void *self)
    {
        HipManager *h = static_cast<HipManager *>(self);
        if (h->log_) h->log_(m);
    }

    fn_engine *eng_ = nullptr;
    log_fn log_;
    std::vector<std::string> names_;
    std::vector<int> ranges_;
    std::vector<std::unique_ptr<DeviceBuffer>> handles_;
};

}  // namespace fibrenet
