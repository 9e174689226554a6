// Messages, verbosity, fatal errors and wall-clock timers with the formats of
// the reference's mod_io (src/mod_io.F90:92-175) and mod_profiling
// (src/mod_profiling.F90:20-71).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>

#include "ommp_internal.h"

namespace ommp {

namespace {
int g_verbose = 1;  // OMMP_VERBOSE_DEFAULT
FILE *g_out = nullptr;
std::mutex g_io_mutex;
constexpr int kTimeStack = 128;
double g_time_stack[kTimeStack];
int g_time_top = 0;

FILE *out() { return g_out ? g_out : stdout; }

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}
}  // namespace

int verbosity() { return g_verbose; }

void message(const std::string &msg, int level, const char *prefix) {
    if (level > g_verbose) return;
    std::string pre;
    if (prefix) {
        pre = std::string("[") + prefix + "]";
    } else if (level == 1) {
        pre = "[warning]";
    } else if (level == 2) {
        pre = "[info]";
    } else if (level == 3) {
        pre = "[debug]";
    }
    std::lock_guard<std::mutex> lock(g_io_mutex);
    // '(A6, A12, " ", A)': the prefix is right-justified in 12 columns
    std::fprintf(out(), "[OMMP]%12s %s\n", pre.substr(0, 12).c_str(), msg.c_str());
    std::fflush(out());
}

namespace {
thread_local bool t_fatal_throws = false;
}

FatalToException::FatalToException() : prev_(t_fatal_throws) { t_fatal_throws = true; }
FatalToException::~FatalToException() { t_fatal_throws = prev_; }

void fatal(const std::string &msg) {
    if (t_fatal_throws) throw FatalError{msg};
    int saved = g_verbose;
    if (g_verbose < 1) g_verbose = 1;
    message(msg, 1, "stop");
    message("Unrecoverable error in openMMPol library. Exiting.", 1, "stop");
    g_verbose = saved;
    if (g_out) {
        std::fclose(g_out);
        g_out = nullptr;
    }
    std::exit(1);
}

}  // namespace ommp

extern "C" {

void ommp_set_verbose(int32_t level) {
    ommp::g_verbose = level < 0 ? 0 : (level > 3 ? 3 : level);
}

void ommp_set_outputfile(const char *filename) {
    if (ommp::g_out) std::fclose(ommp::g_out);
    ommp::g_out = std::fopen(filename, "w");
    if (!ommp::g_out) ommp::fatal(std::string("Cannot open output file ") + filename);
}

void ommp_close_outputfile(void) {
    if (ommp::g_out) std::fclose(ommp::g_out);
    ommp::g_out = nullptr;
}

void ommp_message(const char *msg, int32_t level, const char *prefix) {
    ommp::message(msg ? msg : "", level, (prefix && prefix[0]) ? prefix : nullptr);
}

void ommp_fatal(const char *msg) { ommp::fatal(msg ? msg : ""); }

void ommp_time_push(void) {
    if (ommp::g_time_top >= ommp::kTimeStack) ommp::fatal("Timing stack overflow");
    ommp::g_time_stack[ommp::g_time_top++] = ommp::now_s();
}

void ommp_time_pull(const char *label) {
    if (ommp::g_time_top <= 0) ommp::fatal("Timing stack underflow");
    double t = ommp::now_s() - ommp::g_time_stack[--ommp::g_time_top];
    char buf[4300];
    std::snprintf(buf, sizeof(buf), "%s: %f s", label ? label : "", t);
    ommp::message(std::string("--> ") + buf, 2, "time");
}

}  // extern "C"
