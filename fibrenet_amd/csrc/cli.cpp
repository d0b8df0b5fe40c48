// cli.cpp - fibrenet_run, the command-line driver of a FibreNet simulation on
// the HIP engine (FibreNet.cpp:6-15 without its prompts).
//
//   fibrenet_run [--device N] [--precision f32|f64] [--cadence fast|reference]
//                [--log-dir DIR | --no-log-file] input.txt
//
// The reference asks on stdin for an OpenCL platform, a device
// (OpenclManager.cpp:5-85) and the input file (SimulationManager.cpp:43-44);
// here they are arguments.  It then runs SimulationManager::Initialize, Run and
// Close.  Messages go to stdout and, as log_stream does (log_stream.h:10-67), to
// a log file named log_<sec>_<min>_<hour>_<day>_<month>_<year>.txt (UTC) that is
// opened for appending and starts with "Starting log file at:<date>".
//
// Exit status: 0 after a complete run, 1 when loading, the device or the run
// fails, 2 for a usage error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>

#include "fibrenet.h"

namespace {

struct Logger {
    std::ofstream file;
    void line(const char *m)
    {
        std::cout << m << std::endl;
        if (file.is_open()) file << m << std::endl;
    }
    static void sink(const char *m, void *self) { static_cast<Logger *>(self)->line(m); }
};

int usage(const char *why)
{
    if (why) fprintf(stderr, "fibrenet_run: %s\n", why);
    fprintf(stderr,
            "usage: fibrenet_run [--device N] [--precision f32|f64] [--cadence fast|reference]\n"
            "                    [--log-dir DIR | --no-log-file] input.txt\n"
            "  --device N      HIP device (default 0)\n"
            "  --precision     f32: fp32 forces and velocities, fp64 positions (default); f64: fp64 throughout\n"
            "  --cadence       fast: fused graph-captured steps (default); reference: one launch per kernel\n"
            "                  and a device finish every step, as RunSimulationLoop issues them (same results)\n"
            "  --log-dir DIR   directory of the log file (default: the working directory)\n"
            "  --no-log-file   messages to stdout only\n");
    return 2;
}

bool parse_int(const char *s, int &out)
{
    char *end = nullptr;
    const long v = strtol(s, &end, 10);
    if (!s[0] || *end || v < 0 || v > 1 << 20) return false;
    out = (int)v;
    return true;
}

}  // namespace

int main(int argc, char **argv)
{
    int device = 0, precision = FN_PREC_F32, fast = 1;
    bool log_file = true;
    std::string log_dir, input;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](const char *name) -> const char * {
            if (i + 1 >= argc) return nullptr;
            (void)name;
            return argv[++i];
        };
        if (a == "--device") {
            const char *v = value("--device");
            if (!v || !parse_int(v, device)) return usage("--device needs a non-negative integer");
        } else if (a == "--precision") {
            const char *v = value("--precision");
            if (!v) return usage("--precision needs f32 or f64");
            if (!strcmp(v, "f32")) precision = FN_PREC_F32;
            else if (!strcmp(v, "f64")) precision = FN_PREC_F64;
            else return usage("--precision needs f32 or f64");
        } else if (a == "--cadence") {
            const char *v = value("--cadence");
            if (!v) return usage("--cadence needs fast or reference");
            if (!strcmp(v, "fast")) fast = 1;
            else if (!strcmp(v, "reference")) fast = 0;
            else return usage("--cadence needs fast or reference");
        } else if (a == "--log-dir") {
            const char *v = value("--log-dir");
            if (!v) return usage("--log-dir needs a directory");
            log_dir = v;
        } else if (a == "--no-log-file") {
            log_file = false;
        } else if (a == "-h" || a == "--help") {
            usage(nullptr);
            return 0;
        } else if (!a.empty() && a[0] == '-') {
            return usage(("unknown option " + a).c_str());
        } else if (input.empty()) {
            input = a;
        } else {
            return usage("one input file expected");
        }
    }
    if (input.empty()) return usage("no input file");

    Logger log;
    if (log_file) {
        const time_t now = time(nullptr);
        char date[128] = "";
        struct tm t;
        if (now != (time_t)-1 && gmtime_r(&now, &t)) strftime(date, sizeof date, "%S_%M_%H_%d_%m_%Y", &t);
        const std::string name = (log_dir.empty() ? std::string() : log_dir + "/") + "log_" + date + ".txt";
        log.file.open(name, std::ios_base::app);
        if (log.file.is_open()) log.file << "Starting log file at:" << date << std::endl;
        else fprintf(stderr, "fibrenet_run: unable to open the log file %s; logging to stdout only\n", name.c_str());
    }
    log.line("\tSOLUTION MANAGER INITIALIZATION");
    log.line(("\tInput file: " + input).c_str());
    fn_sim *sim = nullptr;
    int rc = fn_sim_create(device, precision, input.c_str(), &sim);
    if (rc != FN_OK) {
        log.line((std::string("\tError occured during initialization: ") + fn_model_error()).c_str());
        log.line("\tClosing program!");
        return 1;
    }
    fn_sim_set_log_callback(sim, &Logger::sink, &log);   // the engine's messages come through it too
    rc = fn_sim_run(sim, fast);
    if (rc != FN_OK) log.line((std::string("\tError: ") + fn_sim_error(sim)).c_str());
    log.line("\tClosing program.");
    fn_sim_destroy(sim);
    return rc == FN_OK ? 0 : 1;
}
