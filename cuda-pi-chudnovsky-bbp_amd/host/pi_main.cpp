// ============================================================================
//  pi_main.cpp -- `pi_hip`, the reference CLI (src/main.cpp) on the ntt-hip
//  engine: same flags, same output format, `--backend ntt-hip` and
//  `--bbp-device gpu`. There is no CPU path in this build: `--bbp-device cpu`
//  and the reference's CPU backends are rejected with a message instead of
//  being emulated.
// ============================================================================
#include "pidigits_hip.h"
#include "pdh_capi.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

using namespace pidigits;

// The first 1000 decimal and 64 hexadecimal fractional digits of pi, for
// --verify (values of pi computed independently with mpmath).
static const char* PI_DEC_1000 =
    "1415926535897932384626433832795028841971693993751058209749445923078164062862089986280348253421170679"
    "8214808651328230664709384460955058223172535940812848111745028410270193852110555964462294895493038196"
    "4428810975665933446128475648233786783165271201909145648566923460348610454326648213393607260249141273"
    "7245870066063155881748815209209628292540917153643678925903600113305305488204665213841469519415116094"
    "3305727036575959195309218611738193261179310511854807446237996274956735188575272489122793818301194912"
    "9833673362440656643086021394946395224737190702179860943702770539217176293176752384674818467669405132"
    "0005681271452635608277857713427577896091736371787214684409012249534301465495853710507922796892589235"
    "4201995611212902196086403441815981362977477130996051870721134999999837297804995105973173281609631859"
    "5024459455346908302642522308253344685035261931188171010003137838752886587533208381420617177669147303"
    "5982534904287554687311595628638823537875937519577818577805321712268066130019278766111959092164201989";
static const char* PI_HEX_64 = "243F6A8885A308D313198A2E03707344A4093822299F31D0082EFA98EC4E6C89";

static void usage(const char* prog) {
    std::printf(
"pi_hip -- digits of pi on an MI355X (Chudnovsky decimal, or BBP hexadecimal)\n\n"
"  --algo <chudnovsky|bbp>   algorithm (default chudnovsky)\n"
"  --backend <ntt-hip|auto>  multiply backend (every multiply on the device)\n"
"  --digits <N>              decimal digits after the point (default 1000)\n"
"  --bbp-start <P>           1-based hex position (default 1)\n"
"  --bbp-count <C>           hex digits to extract (default 32)\n"
"  --bbp-device <gpu>        where BBP runs (gpu only in this build)\n"
"  --out <file>              write the digits to <file>\n"
"  --verify                  check against pi's first 1000 decimal / 64 hex digits\n"
"  --quiet                   print only the digits\n"
"e.g. %s --digits 1000000 --out pi.txt ; %s --algo bbp --bbp-start 1000000 --bbp-count 16\n", prog, prog);
}

static void emit(const std::string& path, const std::string& data, bool quiet) {
    if (!path.empty()) {
        std::ofstream f(path, std::ios::binary);
        f.write(data.data(), (std::streamsize)data.size());
        f.put('\n');
        if (!quiet) std::printf("wrote %zu characters to %s\n", data.size(), path.c_str());
    } else {
        std::fwrite(data.data(), 1, data.size(), stdout);
        std::putchar('\n');
    }
}

int main(int argc, char** argv) {
    std::string algo = "chudnovsky", backend = "ntt-hip", bbp_device = "gpu", out_path;
    uint64_t digits = 1000, bbp_start = 1;
    uint32_t bbp_count = 32;
    bool verify = false, quiet = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto need = [&](const char* what) -> std::string {
            if (i + 1 >= argc) { std::fprintf(stderr, "error: %s needs a value\n", what); std::exit(2); }
            return argv[++i];
        };
        if (a == "--help" || a == "-h") { usage(argv[0]); return 0; }
        else if (a == "--algo") algo = need("--algo");
        else if (a == "--digits") digits = std::strtoull(need("--digits").c_str(), nullptr, 10);
        else if (a == "--backend") backend = need("--backend");
        else if (a == "--bbp-start") bbp_start = std::strtoull(need("--bbp-start").c_str(), nullptr, 10);
        else if (a == "--bbp-count") bbp_count = (uint32_t)std::strtoul(need("--bbp-count").c_str(), nullptr, 10);
        else if (a == "--bbp-device") bbp_device = need("--bbp-device");
        else if (a == "--out") out_path = need("--out");
        else if (a == "--verify") verify = true;
        else if (a == "--quiet") quiet = true;
        else { std::fprintf(stderr, "error: unknown option '%s' (try --help)\n", a.c_str()); return 2; }
    }
    if (backend != "ntt-hip" && backend != "auto") {
        std::fprintf(stderr, "error: backend '%s' is a CPU backend of the reference; this build has ntt-hip only\n", backend.c_str());
        return 2;
    }
    if (!hip_available()) { std::fprintf(stderr, "error: %s\n", pdh_last_error()); return 1; }
    try {
        if (algo == "chudnovsky") {
            if (!quiet) std::printf("Chudnovsky series + binary splitting\n  digits : %llu\n  backend: ntt-hip\n", (unsigned long long)digits);
            double ph[6] = {0};
            auto t0 = std::chrono::steady_clock::now();
            std::string pi = compute_pi_chudnovsky_hip(digits, ph);
            double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (!quiet) {
                std::printf("  terms  : %llu\n  time   : %.3f s\n", (unsigned long long)ph[5], sec);
                std::printf("  phases : split %.1f ms, 10^prec %.1f ms, sqrt %.1f ms, divide %.1f ms, format %.1f ms\n",
                            ph[0], ph[1], ph[2], ph[3], ph[4]);
            }
            if (verify) {
                std::string frac = pi.size() > 2 ? pi.substr(2) : std::string();
                size_t n = frac.size() < 1000 ? frac.size() : 1000;
                bool good = frac.compare(0, n, std::string(PI_DEC_1000, n)) == 0;
                std::printf("  verify : %s (checked %zu digits)\n", good ? "PASS" : "FAIL", n);
            }
            emit(out_path, pi, quiet);
        } else if (algo == "bbp") {
            if (bbp_device != "gpu") { std::fprintf(stderr, "error: this build runs BBP on the gpu only\n"); return 2; }
            if (!quiet) std::printf("BBP hexadecimal digit extraction\n  start  : %llu (1-based)\n  count  : %u\n  device : gpu\n",
                                    (unsigned long long)bbp_start, bbp_count);
            std::string hex = bbp_hex_digits_hip(bbp_start, bbp_count);
            if (verify && bbp_start >= 1 && bbp_start + bbp_count - 1 <= 64) {
                std::string ref(PI_HEX_64 + (bbp_start - 1), bbp_count);
                std::printf("  verify : %s\n", hex == ref ? "PASS" : "FAIL");
            }
            emit(out_path, hex, quiet);
        } else {
            std::fprintf(stderr, "error: unknown --algo '%s' (use chudnovsky or bbp)\n", algo.c_str());
            return 2;
        }
    } catch (const HipError& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
