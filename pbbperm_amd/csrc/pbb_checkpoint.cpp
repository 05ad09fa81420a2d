// pbb_checkpoint.cpp -- the work-unit codec of the reference's distributed
// layer, without GMP: factoradic digit vectors <-> decimal interval ranks
// (fact_work::VectToBigint / BigintToVect, distributed/src/fact_work.cpp:47-86,
// with the weights of distributed/src/weights.cpp:10-26) and the checkpoint
// text of works::save (distributed/src/works.cpp:172-191), whose pieces are
// solution operator<< (common/src/solution.cpp:132-142), works operator<<
// (distributed/src/works.cpp:558-569) and work::writeHeader /
// writeIntervals (distributed/src/work.cpp:544-583).
//
// Ranks are at most 63! - 1 < 2^297, held in six 64-bit limbs.
#include <cctype>
#include <climits>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pbb_engine.h"
#include "pbb_util.h"

namespace {

constexpr int LIMBS = 6;

struct Rank {
    uint64_t w[LIMBS] = {0, 0, 0, 0, 0, 0};
};

bool is_zero(const Rank &r)
{
    for (uint64_t x : r.w)
        if (x) return false;
    return true;
}

// r = r * m + a; false on overflow
bool mul_add(Rank &r, uint32_t m, uint32_t a)
{
    unsigned __int128 carry = a;
    for (int i = 0; i < LIMBS; i++) {
        unsigned __int128 t = (unsigned __int128)r.w[i] * m + carry;
        r.w[i] = (uint64_t)t;
        carry = t >> 64;
    }
    return carry == 0;
}

// r = r / d, returns r mod d
uint32_t div_small(Rank &r, uint32_t d)
{
    unsigned __int128 rem = 0;
    for (int i = LIMBS - 1; i >= 0; i--) {
        unsigned __int128 cur = (rem << 64) | r.w[i];
        r.w[i] = (uint64_t)(cur / d);
        rem = cur % d;
    }
    return (uint32_t)rem;
}

int cmp(const Rank &a, const Rank &b)
{
    for (int i = LIMBS - 1; i >= 0; i--)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
}

// a - b for a >= b
Rank sub(const Rank &a, const Rank &b)
{
    Rank r;
    unsigned borrow = 0;
    for (int i = 0; i < LIMBS; i++) {
        uint64_t x = a.w[i], y = b.w[i];
        uint64_t d = x - y - borrow;
        borrow = (x < y || (x == y && borrow)) ? 1 : 0;
        r.w[i] = d;
    }
    return r;
}

Rank add(const Rank &a, const Rank &b)
{
    Rank r;
    unsigned __int128 c = 0;
    for (int i = 0; i < LIMBS; i++) {
        unsigned __int128 t = (unsigned __int128)a.w[i] + b.w[i] + c;
        r.w[i] = (uint64_t)t;
        c = t >> 64;
    }
    return r;
}

std::string to_dec(Rank r)
{
    if (is_zero(r)) return "0";
    std::string s;
    while (!is_zero(r)) s.push_back((char)('0' + div_small(r, 10)));
    return std::string(s.rbegin(), s.rend());
}

bool from_dec(const std::string &s, Rank &r)
{
    r = Rank();
    if (s.empty()) return false;
    for (char ch : s) {
        if (!isdigit((unsigned char)ch)) return false;
        if (!mul_add(r, 10, (uint32_t)(ch - '0'))) return false;
    }
    return true;
}

// sum of digits[i] * (n-1-i)!  (Horner over the mixed radix n, n-1, ..., 1)
bool digits_to_rank(int n, const int *d, Rank &r)
{
    r = Rank();
    for (int i = 0; i < n; i++) {
        if (d[i] < 0 || d[i] > n - 1 - i) return false;
        if (!mul_add(r, (uint32_t)(i == 0 ? 1 : n - i), (uint32_t)d[i])) return false;
    }
    return true;
}

bool rank_to_digits(int n, Rank r, int *d)
{
    for (int i = n - 1; i >= 1; i--) d[i] = (int)div_small(r, (uint32_t)(n - i));
    // what is left is the leading digit, below n
    Rank lim;
    mul_add(lim, 1, (uint32_t)n);
    if (cmp(r, lim) >= 0) return false;
    d[0] = (int)r.w[0];
    return true;
}

void write_solution(std::ostringstream &o, int n, int cost, const int *perm)
{
    o << n << "\n" << cost << "\n";
    for (int i = 0; i < n; i++) o << perm[i] << " ";
    o << "\n";
}

bool read_solution(std::istringstream &in, int n, int *cost, int *perm)
{
    int size = 0;
    if (!(in >> size) || size != n) return false;
    if (!(in >> *cost)) return false;
    for (int i = 0; i < n; i++)
        if (!(in >> perm[i])) return false;
    return true;
}

}  // namespace

using pbb::fail;

extern "C" int pbb_fact_to_dec(int n, const int *digits, char *dec, int cap)
{
    if (n < 1 || n > PBB_MAX_JOBS || !digits || !dec || cap < 2) return fail(PBB_E_ARG, "bad arguments");
    Rank r;
    if (!digits_to_rank(n, digits, r)) return fail(PBB_E_ARG, "digit out of its factoradic range");
    std::string s = to_dec(r);
    if ((int)s.size() + 1 > cap) return fail(PBB_E_CAPACITY, "decimal buffer too small");
    memcpy(dec, s.c_str(), s.size() + 1);
    return PBB_OK;
}

extern "C" int pbb_dec_to_fact(int n, const char *dec, int *digits)
{
    if (n < 1 || n > PBB_MAX_JOBS || !dec || !digits) return fail(PBB_E_ARG, "bad arguments");
    Rank r;
    if (!from_dec(dec, r)) return fail(PBB_E_ARG, "'%s' is not a decimal rank", dec);
    if (!rank_to_digits(n, r, digits)) return fail(PBB_E_ARG, "rank %s is not below %d!", dec, n);
    return PBB_OK;
}

extern "C" int pbb_checkpoint_format(const pbb_checkpoint *c, const int *ids, const int *pos, const int *end,
                                     char *buf, int cap, int *len)
{
    if (!c || c->n < 1 || c->n > PBB_MAX_JOBS || c->nb < 0 || (c->nb > 0 && (!ids || !pos || !end)) || !len)
        return fail(PBB_E_ARG, "bad checkpoint arguments");
    const int n = c->n;
    std::ostringstream o;
    write_solution(o, n, c->root_cost, c->root_perm);
    o << c->simple_bounds << " ";
    write_solution(o, n, c->best_cost, c->best_perm);
    // works: one unit holding every interval
    o << 1 << "\n";
    o << c->work_id << " " << 0 << " " << c->max_intervals << " " << c->explored << " " << c->nb << "\n";
    Rank total;
    for (int k = 0; k < c->nb; k++) {
        Rank b, e;
        if (!digits_to_rank(n, pos + (size_t)k * n, b) || !digits_to_rank(n, end + (size_t)k * n, e))
            return fail(PBB_E_ARG, "interval %d: digit out of its factoradic range", k);
        if (cmp(e, b) > 0) total = add(total, sub(e, b));
        o << ids[k] << " " << to_dec(b) << " " << to_dec(e) << "\n";
    }
    o << to_dec(total) << "\n";
    const std::string s = o.str();
    *len = (int)s.size();
    if (!buf) return PBB_OK;  // size query
    if ((int)s.size() + 1 > cap) return fail(PBB_E_CAPACITY, "checkpoint buffer too small (%d bytes needed)",
                                             (int)s.size() + 1);
    memcpy(buf, s.c_str(), s.size() + 1);
    return PBB_OK;
}

extern "C" int pbb_checkpoint_parse(const char *text, int n, pbb_checkpoint *c, int *ids, int *pos, int *end,
                                    int max)
{
    if (!text || !c || n < 1 || n > PBB_MAX_JOBS || max < 0 || (max > 0 && (!ids || !pos || !end)))
        return fail(PBB_E_ARG, "bad checkpoint arguments");
    memset(c, 0, sizeof *c);
    c->n = n;
    std::istringstream in(text);
    if (!read_solution(in, n, &c->root_cost, c->root_perm)) return fail(PBB_E_ARG, "checkpoint: root solution");
    unsigned long long sb = 0;
    if (!(in >> sb)) return fail(PBB_E_ARG, "checkpoint: simpleBounds");
    c->simple_bounds = sb;
    if (!read_solution(in, n, &c->best_cost, c->best_perm)) return fail(PBB_E_ARG, "checkpoint: incumbent");
    int units = 0;
    if (!(in >> units) || units < 0) return fail(PBB_E_ARG, "checkpoint: number of work units");
    int nb = 0;
    for (int u = 0; u < units; u++) {
        // work::readHeader: id end_updated max_intervals exploredNodes nb_intervals
        long long wid = 0, upd = 0, mx = 0;
        std::string explored;
        long long cnt = 0;
        if (!(in >> wid >> upd >> mx >> explored >> cnt) || cnt < 0)
            return fail(PBB_E_ARG, "checkpoint: header of work unit %d", u);
        Rank ex;
        if (!from_dec(explored, ex)) return fail(PBB_E_ARG, "checkpoint: exploredNodes of unit %d", u);
        c->explored += ex.w[0];
        if (u == 0) {
            c->work_id = (int)wid;
            c->max_intervals = (int)mx;
        }
        for (long long k = 0; k < cnt; k++) {
            std::string iid, bb, ee;
            if (!(in >> iid >> bb >> ee)) return fail(PBB_E_ARG, "checkpoint: interval %lld of unit %d", k, u);
            Rank b, e;
            if (!from_dec(bb, b) || !from_dec(ee, e)) return fail(PBB_E_ARG, "checkpoint: interval bounds");
            if (cmp(b, e) > 0) continue;  // empty
            if (nb >= max) return fail(PBB_E_CAPACITY, "checkpoint holds more than %d intervals", max);
            if (!rank_to_digits(n, b, pos + (size_t)nb * n) || !rank_to_digits(n, e, end + (size_t)nb * n))
                return fail(PBB_E_ARG, "checkpoint: interval rank not below %d!", n);
            ids[nb] = atoi(iid.c_str());
            nb++;
        }
        std::string size;
        if (!(in >> size)) return fail(PBB_E_ARG, "checkpoint: size of work unit %d", u);
    }
    c->nb = nb;
    return PBB_OK;
}
