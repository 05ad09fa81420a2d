// fuzz_parsers.cpp -- seeded mutation fuzzing of the library's text and
// file parsers, built with the CPU AddressSanitizer and UndefinedBehavior-
// Sanitizer (tests/test_sanitize.py compiles it with g++ together with
// pbb_io.cpp and pbb_checkpoint.cpp, the device-free host sources).
//
//   fuzz_parsers <seed-dir> <work-dir> <iterations per seed file>
//
// Seed files are named for the parser they feed: ini_*, vrf_*, ckpt_<n>_*
// and dec_<n>_*.  Each is parsed as it is, then mutated (byte flips,
// deletions, insertions of digits, signs, separators and long numbers,
// truncation, duplication of a slice) and parsed again, with every output
// buffer allocated at exactly the size the call is given, so a write past it
// is a sanitizer report.  Codec round trips run on random digit vectors and
// on every buffer capacity around the needed one.  A sanitizer report aborts
// the process (-fno-sanitize-recover=all); a round-trip mismatch exits 2.
#include <dirent.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "pbb_engine.h"

namespace {

uint64_t g_state = 0x243F6A8885A308D3ull;

uint64_t next_u64()
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

size_t below(size_t k) { return k ? (size_t)(next_u64() % k) : 0; }

std::string slurp(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

void spit(const std::string &path, const std::string &s)
{
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    f << s;
}

std::string mutate(std::string s)
{
    static const char *tokens[] = {"-", "0", "9", " ", "\n", "=", ";", "[", "]", "#", "\t", "-1",
                                   "2147483647", "4294967296", "99999999999999999999999999999999",
                                   "18446744073709551616", "-2147483648", "1e9", "0x7fffffff", "63", "64"};
    const int edits = 1 + (int)below(4);
    for (int e = 0; e < edits; e++) {
        switch (below(6)) {
        case 0:  // flip one byte
            if (!s.empty()) s[below(s.size())] = (char)below(256);
            break;
        case 1:  // delete a slice
            if (!s.empty()) {
                size_t a = below(s.size());
                s.erase(a, 1 + below(16));
            }
            break;
        case 2: {  // insert a token
            const char *t = tokens[below(sizeof tokens / sizeof *tokens)];
            s.insert(below(s.size() + 1), t);
            break;
        }
        case 3:  // truncate
            s.resize(below(s.size() + 1));
            break;
        case 4:  // duplicate a slice
            if (!s.empty()) {
                size_t a = below(s.size()), l = 1 + below(64);
                std::string piece = s.substr(a, l);
                s.insert(below(s.size() + 1), piece);
            }
            break;
        default:  // replace a digit run by a long number
            for (size_t i = below(s.size() + 1); i < s.size(); i++)
                if (s[i] >= '0' && s[i] <= '9') {
                    size_t j = i;
                    while (j < s.size() && s[j] >= '0' && s[j] <= '9') j++;
                    s.replace(i, j - i, tokens[12 + below(6)]);
                    break;
                }
            break;
        }
    }
    return s;
}

int parse_n(const std::string &name, const char *prefix)
{
    // <prefix><n>_...
    return atoi(name.c_str() + strlen(prefix));
}

void run_ini(const std::string &text, const std::string &tmp)
{
    spit(tmp, text);
    pbb_config c;
    std::unique_ptr<char[]> problem(new char[64]), instance(new char[64]);
    (void)pbb_config_load_ini(tmp.c_str(), &c, problem.get(), instance.get());
}

void run_vrf(const std::string &text, const std::string &tmp)
{
    spit(tmp, text);
    for (int cap : {0, 1, 7, 100, 1260}) {
        std::unique_ptr<int32_t[]> ptm(new int32_t[cap ? cap : 1]);
        int n = 0, m = 0;
        (void)pbb_instance_read_vrf(tmp.c_str(), &n, &m, ptm.get(), cap);
    }
}

void run_ckpt(const std::string &text, int n)
{
    for (int max : {0, 1, 3, 64}) {
        std::unique_ptr<int[]> ids(new int[max ? max : 1]);
        std::unique_ptr<int[]> pos(new int[(size_t)(max ? max : 1) * n]), end(new int[(size_t)(max ? max : 1) * n]);
        pbb_checkpoint c;
        int rc = pbb_checkpoint_parse(text.c_str(), n, &c, max ? ids.get() : nullptr, max ? pos.get() : nullptr,
                                      max ? end.get() : nullptr, max);
        if (rc == PBB_OK && c.nb > 0) {
            // a parsed text formats again, at every capacity around the size
            int len = 0;
            if (pbb_checkpoint_format(&c, ids.get(), pos.get(), end.get(), nullptr, 0, &len) != PBB_OK) exit(2);
            for (int cap = len - 2; cap <= len + 1; cap++) {
                if (cap < 1) continue;
                std::unique_ptr<char[]> buf(new char[cap]);
                int l2 = 0;
                int r2 = pbb_checkpoint_format(&c, ids.get(), pos.get(), end.get(), buf.get(), cap, &l2);
                if ((r2 == PBB_OK) != (cap >= len + 1)) exit(2);
            }
        }
    }
}

void run_dec(const std::string &text, int n)
{
    std::unique_ptr<int[]> d(new int[n]);
    std::string z = text;  // parser takes a C string
    if (pbb_dec_to_fact(n, z.c_str(), d.get()) == PBB_OK) {
        for (int cap : {2, 8, 40, 100}) {
            std::unique_ptr<char[]> buf(new char[cap]);
            int rc = pbb_fact_to_dec(n, d.get(), buf.get(), cap);
            if (rc == PBB_OK) {
                std::unique_ptr<int[]> d2(new int[n]);
                if (pbb_dec_to_fact(n, buf.get(), d2.get()) != PBB_OK) exit(2);
                if (memcmp(d.get(), d2.get(), sizeof(int) * n) != 0) exit(2);
            }
        }
    }
}

// random digit vectors: fact -> dec -> fact at every capacity up to the size
void codec_round_trips(int count)
{
    for (int it = 0; it < count; it++) {
        const int n = 1 + (int)below(PBB_MAX_JOBS);
        std::unique_ptr<int[]> d(new int[n]), back(new int[n]);
        for (int i = 0; i < n; i++) d[i] = (int)below((size_t)(n - i));
        if (below(8) == 0) d[below(n)] = n + 1;  // out of range: must be refused
        char big[128];
        int rc = pbb_fact_to_dec(n, d.get(), big, sizeof big);
        bool valid = true;
        for (int i = 0; i < n; i++) valid = valid && d[i] >= 0 && d[i] < n - i;
        if ((rc == PBB_OK) != valid) exit(2);
        if (!valid) continue;
        const int need = (int)strlen(big) + 1;
        for (int cap = 2; cap <= need; cap++) {
            std::unique_ptr<char[]> buf(new char[cap]);
            int r2 = pbb_fact_to_dec(n, d.get(), buf.get(), cap);
            if ((r2 == PBB_OK) != (cap >= need)) exit(2);
        }
        if (pbb_dec_to_fact(n, big, back.get()) != PBB_OK) exit(2);
        if (memcmp(d.get(), back.get(), sizeof(int) * n) != 0) exit(2);
    }
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s <seed-dir> <work-dir> <iterations>\n", argv[0]);
        return 64;
    }
    const std::string dir = argv[1], work = argv[2];
    const int iters = atoi(argv[3]);
    std::vector<std::string> names;
    if (DIR *d = opendir(dir.c_str())) {
        while (dirent *ent = readdir(d))
            if (ent->d_name[0] != '.') names.push_back(ent->d_name);
        closedir(d);
    }
    const std::string tmp = work + "/fuzz_input.txt";
    long runs = 0;
    for (const std::string &name : names) {
        const std::string seed = slurp(dir + "/" + name);
        for (int it = 0; it <= iters; it++) {
            const std::string text = it == 0 ? seed : mutate(seed);
            if (name.rfind("ini_", 0) == 0) run_ini(text, tmp);
            else if (name.rfind("vrf_", 0) == 0) run_vrf(text, tmp);
            else if (name.rfind("ckpt_", 0) == 0) run_ckpt(text, parse_n(name, "ckpt_"));
            else if (name.rfind("dec_", 0) == 0) run_dec(text, parse_n(name, "dec_"));
            else continue;
            runs++;
        }
    }
    codec_round_trips(20000);
    printf("fuzz_parsers: %zu seed files, %ld parser runs, 20000 codec round trips\n", names.size(), runs);
    return 0;
}
