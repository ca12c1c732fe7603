// ============================================================================
//  oracle/seam_probe.cpp -- TEST INFRASTRUCTURE ONLY.
//
//  Linked into the reference programs that oracle/Makefile's `seam` target
//  builds from the reference's unchanged sources against libpdh_seam.so. At
//  exit it prints the number of kernels libpdh.so launched in the process,
//  so a test can tell the device path from a host fallback (the reference's
//  own self-test cannot, tests/self_test.cu:66).
// ============================================================================
#include "pdh_capi.h"
#include "pdh_diag.h"

#include <cstdio>

namespace {
struct LaunchReport {
    ~LaunchReport() {
        std::fprintf(stderr, "pdh_launch_count=%llu\n", (unsigned long long)pdh_launch_count());
        std::fflush(stderr);
    }
} g_report;
} // namespace
