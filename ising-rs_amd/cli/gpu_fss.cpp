// gpu_fss: finite-size-scaling parallel tempering for Ising, XY and Heisenberg
// on one MI355X.  Drop-in for the reference's src/bin/gpu_fss.rs: same flags,
// defaults, validation messages and output files:
//
//   gpu_fss --model ising --sizes 8,16,32,64 --tmin 4.4 --tmax 4.6 --replicas 20
//           --warmup 5000 --samples 100000 --exchange-every 10 --seed 42
//           --outdir analysis/data [--delta 0.5] [--overrelax 5] [--measure-every 1]
//
// Output per size (gpu_fss.rs:11-13):
//   gpu_fss_{model}_N{n}_summary.csv     T,E,E_err,M,M_err,M2,M2_err,M4,M4_err,Cv,Cv_err,chi,chi_err
//   gpu_fss_{model}_N{n}_timeseries.csv  temp_idx,E,M (one row per replica per measured sweep)
//
// Additional flags (off by default, results identical either way):
//   --batched 0|1      run every size in one ensemble (1, default) or one size at a time (0)
//   --renorm-every k   continuous spins: renormalise every k sweeps (default 100)
//   --binary 0|1       also write gpu_fss_{model}_N{n}_timeseries.bin: the exact
//                      per-replica records (int64 bond sum and magnetisation for
//                      Ising, FP64 E, Mx, My, Mz otherwise), which the 8-decimal CSV
//                      rounds; format below
//
// Binary timeseries (little-endian), one file per size:
//   char[8] "ISINGTS1"; int32 version (1), model (0 ising, 1 xy, 2 heisenberg),
//   L, replicas R, measure_every, record kind (0: Ising, 1: continuous);
//   int64 number of measured sweeps; double J; double T[R] (by temperature
//   index); then, per measured sweep, R records in replica order:
//     kind 0: int64 bond sum (E = -J * bond), int64 magnetisation, int32 temp_idx, int32 0
//     kind 1: double E, Mx, My, Mz (totals), int32 temp_idx, int32 0
//
// Divergences from the reference, by design: initial states and exchange
// draws come from the library's Philox stream (the reference's are unseeded /
// Xoshiro256++); --exchange-every 0 is rejected instead of panicking.
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ising_hip.h"

namespace {

[[noreturn]] void die(const std::string& msg) {
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(1);
}

std::string get_arg(int argc, char** argv, int i, const char* flag) {  // cli.rs:10-16
    if (i + 1 >= argc) die(std::string("Error: ") + flag + " requires a value");
    return argv[i + 1];
}

// Parse errors worded like Rust's FromStr messages (cli.rs:19-28).
unsigned long long parse_u64(const std::string& v, const char* flag) {
    const char* why = nullptr;
    const size_t first = (!v.empty() && v[0] == '+') ? 1 : 0;
    if (v.empty()) why = "cannot parse integer from empty string";
    else if (first == v.size()) why = "invalid digit found in string";
    for (size_t k = first; !why && k < v.size(); ++k)
        if (v[k] < '0' || v[k] > '9') why = "invalid digit found in string";
    unsigned long long out = 0;
    if (!why) {
        errno = 0;
        out = std::strtoull(v.c_str() + first, nullptr, 10);
        if (errno == ERANGE) why = "number too large to fit in target type";
    }
    if (why) die("Error: invalid value '" + v + "' for " + flag + ": " + why);
    return out;
}

double parse_f64(const std::string& v, const char* flag) {
    char* end = nullptr;
    errno = 0;
    const double d = v.empty() ? 0.0 : std::strtod(v.c_str(), &end);
    bool ok = !v.empty() && end && *end == '\0' && !std::isspace((unsigned char)v[0]);
    if (ok && v.size() >= 2 && (v[0] == '0' || v[1] == '0') && (v.find('x') != std::string::npos ||
                                                                v.find('X') != std::string::npos))
        ok = false;  // Rust rejects hex floats
    if (!ok) die("Error: invalid value '" + v + "' for " + flag + ": invalid float literal");
    return d;
}

// Shortest decimal that reads back to the same double, as Rust's Display prints it.
std::string show(double x) {
    char buf[64];
    for (int p = 1; p <= 17; ++p) {
        std::snprintf(buf, sizeof buf, "%.*g", p, x);
        if (std::strtod(buf, nullptr) == x) break;
    }
    std::string s = buf;
    if (s.find('e') != std::string::npos) {
        std::snprintf(buf, sizeof buf, "%.17f", x);
        s = buf;
        while (!s.empty() && s.back() == '0') s.pop_back();
        if (!s.empty() && s.back() == '.') s.pop_back();
    }
    return s;
}

void validate_samples(unsigned long long n, const char* flag) {  // cli.rs:86-91
    if (n == 0) die(std::string("Error: ") + flag + " must be at least 1");
}

void validate_temp_range(double t_min, double t_max) {  // cli.rs:94-107
    if (t_min <= 0.0) die("Error: --tmin must be positive (got " + show(t_min) + ")");
    if (t_max <= 0.0) die("Error: --tmax must be positive (got " + show(t_max) + ")");
    if (t_min >= t_max) die("Error: --tmin (" + show(t_min) + ") must be less than --tmax (" + show(t_max) + ")");
}

void mkdirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && cur != "." && ::mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST)
                die("failed to create outdir: " + std::string(std::strerror(errno)));
        }
        if (i < path.size()) cur.push_back(path[i]);
    }
}

void check(int rc, const char* what) {
    if (rc != ISING_OK) die(std::string("Error: ") + what + ": " + ising_last_error());
}

struct Args {
    std::string model = "ising";
    std::string sizes_str = "8,16,32,64";
    double t_min = 4.4, t_max = 4.6;
    unsigned long long replicas = 20, warmup = 5000, samples = 100000, exchange_every = 10, seed = 42;
    std::string outdir = "analysis/data";
    double delta = 0.5;
    unsigned long long n_overrelax = 5, measure_every = 1;
    unsigned long long batched = 1, renorm_every = 100, binary = 0;
    // --algorithm (the flag the reference plans, gpu-compute-optimizations-impl.md:1348-1380):
    // metropolis; msc (Ising, multispin coded); sw / wolff (a Swendsen-Wang or Wolff cluster
    // update after every Metropolis sweep)
    std::string algorithm = "metropolis";
};

// Writer of the binary timeseries of one size.
struct BinSeries {
    FILE* f = nullptr;
    int64_t n = 0;
    void open(const std::string& path, int model, int L, int R, int measure_every, double J,
              const std::vector<double>& temps) {
        f = std::fopen(path.c_str(), "wb");
        if (!f) die("create timeseries file failed: " + path);
        std::setvbuf(f, nullptr, _IOFBF, 1 << 20);
        const bool ints = model == ISING_MODEL_ISING || model == ISING_MODEL_ISING_MSC;
        const int32_t hdr[6] = {1, model, L, R, measure_every, ints ? 0 : 1};
        const int64_t zero = 0;
        if (std::fwrite("ISINGTS1", 1, 8, f) != 8 || std::fwrite(hdr, sizeof hdr, 1, f) != 1 ||
            std::fwrite(&zero, sizeof zero, 1, f) != 1 || std::fwrite(&J, sizeof J, 1, f) != 1 ||
            std::fwrite(temps.data(), sizeof(double), temps.size(), f) != temps.size())
            die("write timeseries failed: " + path);
    }
    void put(const void* rec, size_t bytes) {
        if (std::fwrite(rec, 1, bytes, f) != bytes) die("write timeseries failed");
    }
    void close(const std::string& path, int64_t sweeps) {
        if (std::fseek(f, 32, SEEK_SET) != 0 || std::fwrite(&sweeps, sizeof sweeps, 1, f) != 1 || std::fclose(f) != 0)
            die("write timeseries failed: " + path);
        f = nullptr;
    }
};

// One FSS run over `sizes` sharing one device ensemble.
void run(const Args& a, const std::vector<int>& sizes, int model, const char* model_name, int cluster) {
    const int R = (int)a.replicas;
    const bool ints = model == ISING_MODEL_ISING || model == ISING_MODEL_ISING_MSC;  // int64 records
    // `n` sweeps of the chosen algorithm: Metropolis sweeps, each followed by a cluster update when asked
    auto sweeps = [&](ising_pt_t* pt, int64_t n, bool sample) {
        if (cluster < 0) return sample ? ising_pt_sample(pt, n) : ising_pt_warmup(pt, n);
        for (int64_t k = 0; k < n; ++k) {
            int rc = sample ? ising_pt_sample(pt, 1) : ising_pt_warmup(pt, 1);
            if (!rc) rc = ints ? ising_pt_cluster(pt, cluster, 1, nullptr, nullptr, nullptr)
                               : ising_pt_cluster_cont(pt, cluster, 1, nullptr, nullptr);
            if (rc) return rc;
        }
        return 0;
    };
    std::vector<double> temps(R);
    check(ising_linspace_temperatures(a.t_min, a.t_max, R, temps.data()), "linspace");
    const int ns = (int)sizes.size();
    for (int L : sizes) std::fprintf(stderr, "  N=%d: creating %d %sreplicas...\n", L, R,
                                     ints ? "" : (std::string(model_name) + " ").c_str());
    ising_pt_opts_t o;
    ising_pt_default_opts(&o);
    o.j = 1.0;
    o.h = 0.0;
    o.exchange_every = (int32_t)a.exchange_every;
    o.measure_every = (int32_t)a.measure_every;
    o.delta = (double)(float)a.delta;
    o.n_overrelax = (int32_t)a.n_overrelax;
    o.renorm_every = (int32_t)a.renorm_every;
    const long long chunk = 1000;
    o.record_capacity = (int32_t)(chunk / (long long)a.measure_every + 2);
    ising_pt_t* pt = nullptr;
    std::vector<int32_t> sz(sizes.begin(), sizes.end());
    check(ising_pt_create(&pt, model, sz.data(), ns, R, temps.data(), a.seed, &o), "failed to create GPU ensemble");

    for (int L : sizes) std::fprintf(stderr, "  N=%d: warming up %llu sweeps...\n", L, a.warmup);
    for (unsigned long long w = 0; w < a.warmup;) {
        const unsigned long long step = std::min<unsigned long long>(1000 - (w % 1000), a.warmup - w);
        check(sweeps(pt, (int64_t)step, false), "GPU step failed");
        w += step;
        if (w % 1000 == 0) {
            check(ising_pt_sync(pt, nullptr), "GPU step failed");
            std::fprintf(stderr, "    warmup %llu/%llu\n", w, a.warmup);
        }
    }

    std::vector<FILE*> ts(ns);
    std::vector<std::string> ts_path(ns), sum_path(ns), bin_path(ns);
    std::vector<BinSeries> bin(a.binary ? ns : 0);
    for (int g = 0; g < (int)bin.size(); ++g) {
        bin_path[g] = a.outdir + "/gpu_fss_" + model_name + "_N" + std::to_string(sizes[g]) + "_timeseries.bin";
        bin[g].open(bin_path[g], model, sizes[g], R, (int)a.measure_every, 1.0, temps);
    }
    int64_t measured = 0;
    for (int g = 0; g < ns; ++g) {
        ts_path[g] = a.outdir + "/gpu_fss_" + model_name + "_N" + std::to_string(sizes[g]) + "_timeseries.csv";
        sum_path[g] = a.outdir + "/gpu_fss_" + model_name + "_N" + std::to_string(sizes[g]) + "_summary.csv";
        ts[g] = std::fopen(ts_path[g].c_str(), "w");
        if (!ts[g]) die("create timeseries file failed: " + ts_path[g]);
        std::setvbuf(ts[g], nullptr, _IOFBF, 1 << 20);
        std::fputs("temp_idx,E,M\n", ts[g]);
    }
    // per (size, temperature): (e_per, m_per) series
    std::vector<std::vector<std::vector<double>>> se(ns, std::vector<std::vector<double>>(R)),
        sm(ns, std::vector<std::vector<double>>(R));
    for (int L : sizes) {
        if (ints)  // gpu_fss.rs:234 / :493
            std::fprintf(stderr, "  N=%d: sampling %llu sweeps with PT exchange every %llu...\n", L, a.samples,
                         a.exchange_every);
        else
            std::fprintf(stderr, "  N=%d: sampling %llu sweeps with PT (measure every %llu)...\n", L, a.samples,
                         a.measure_every);
    }

    const int64_t tot = (int64_t)ns * R;
    const int64_t cap = o.record_capacity;
    std::vector<int32_t> tidx(cap * tot);
    std::vector<int64_t> bond(cap * tot), mag(cap * tot);
    std::vector<double> ce(cap * tot), cx(cap * tot), cy(cap * tot), cz(cap * tot);
    for (unsigned long long s = 0; s < a.samples;) {
        const unsigned long long step = std::min<unsigned long long>(chunk - (s % chunk), a.samples - s);
        check(sweeps(pt, (int64_t)step, true), "GPU step failed");
        s += step;
        int64_t n = 0;
        if (ints)
            check(ising_pt_read_ising(pt, cap, tidx.data(), bond.data(), mag.data(), &n), "GPU measure failed");
        else
            check(ising_pt_read_cont(pt, cap, tidx.data(), ce.data(), cx.data(), cy.data(), cz.data(), &n),
                  "GPU measure failed");
        measured += n;
        for (int64_t k = 0; k < n; ++k)
            for (int g = 0; g < ns; ++g) {
                const double n3 = (double)sizes[g] * sizes[g] * sizes[g];
                for (int r = 0; r < R; ++r) {
                    const int64_t i = k * tot + (int64_t)g * R + r;
                    double e_per, m_per;
                    if (ints) {
                        e_per = -1.0 * (double)bond[i] / n3;  // gpu_fss.rs:250-251
                        m_per = std::fabs((double)mag[i]) / n3;
                    } else {
                        e_per = ce[i] / n3;  // gpu_fss.rs:510-511
                        m_per = std::sqrt(cx[i] * cx[i] + cy[i] * cy[i] + cz[i] * cz[i]) / n3;
                    }
                    const int t = tidx[i];
                    se[g][t].push_back(e_per);
                    sm[g][t].push_back(m_per);
                    std::fprintf(ts[g], "%d,%.8f,%.8f\n", t, e_per, m_per);  // gpu_fss.rs:264
                    if (a.binary) {
                        if (ints) {
                            const int64_t rec[3] = {bond[i], mag[i], (int64_t)(uint32_t)t};
                            bin[g].put(rec, sizeof rec);
                        } else {
                            const double rec[4] = {ce[i], cx[i], cy[i], cz[i]};
                            const int64_t ti = (int64_t)(uint32_t)t;
                            bin[g].put(rec, sizeof rec);
                            bin[g].put(&ti, sizeof ti);
                        }
                    }
                }
            }
        if (s % 10000 == 0) std::fprintf(stderr, "    sweep %llu/%llu\n", s, a.samples);
    }
    for (int g = 0; g < ns; ++g) {
        if (std::fclose(ts[g]) != 0) die("flush timeseries failed");
        std::fprintf(stderr, "  Wrote %s\n", ts_path[g].c_str());
        if (a.binary) {
            bin[g].close(bin_path[g], measured);
            std::fprintf(stderr, "  Wrote %s\n", bin_path[g].c_str());
        }
        const double n3 = (double)sizes[g] * sizes[g] * sizes[g];
        std::string csv = "T,E,E_err,M,M_err,M2,M2_err,M4,M4_err,Cv,Cv_err,chi,chi_err\n";  // gpu_fss.rs:295
        for (int t = 0; t < R; ++t) {
            if (se[g][t].empty()) continue;
            const double T = temps[t], beta = 1.0 / T;
            const int nb = (int)std::min<size_t>(20, se[g][t].size());
            double obs[12];
            check(ising_jackknife(se[g][t].data(), sm[g][t].data(), (int64_t)se[g][t].size(), beta, n3, nb, obs),
                  "jackknife");
            char line[512];
            std::snprintf(line, sizeof line, "%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f\n", T,
                          obs[0], obs[1], obs[2], obs[3], obs[4], obs[5], obs[6], obs[7], obs[8], obs[9], obs[10],
                          obs[11]);
            csv += line;
        }
        FILE* f = std::fopen(sum_path[g].c_str(), "w");
        if (!f || std::fwrite(csv.data(), 1, csv.size(), f) != csv.size() || std::fclose(f) != 0)
            die("write summary failed: " + sum_path[g]);
        std::fprintf(stderr, "  Wrote %s\n", sum_path[g].c_str());
    }
    ising_pt_destroy(pt);
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc;) {  // gpu_fss.rs:37-96; unknown arguments are skipped
        const std::string f = argv[i];
        if (f == "--model") a.model = get_arg(argc, argv, i, "--model"), i += 2;
        else if (f == "--sizes") a.sizes_str = get_arg(argc, argv, i, "--sizes"), i += 2;
        else if (f == "--tmin") a.t_min = parse_f64(get_arg(argc, argv, i, "--tmin"), "--tmin"), i += 2;
        else if (f == "--tmax") a.t_max = parse_f64(get_arg(argc, argv, i, "--tmax"), "--tmax"), i += 2;
        else if (f == "--replicas") a.replicas = parse_u64(get_arg(argc, argv, i, "--replicas"), "--replicas"), i += 2;
        else if (f == "--warmup") a.warmup = parse_u64(get_arg(argc, argv, i, "--warmup"), "--warmup"), i += 2;
        else if (f == "--samples") a.samples = parse_u64(get_arg(argc, argv, i, "--samples"), "--samples"), i += 2;
        else if (f == "--exchange-every")
            a.exchange_every = parse_u64(get_arg(argc, argv, i, "--exchange-every"), "--exchange-every"), i += 2;
        else if (f == "--seed") a.seed = parse_u64(get_arg(argc, argv, i, "--seed"), "--seed"), i += 2;
        else if (f == "--outdir") a.outdir = get_arg(argc, argv, i, "--outdir"), i += 2;
        else if (f == "--delta") a.delta = parse_f64(get_arg(argc, argv, i, "--delta"), "--delta"), i += 2;
        else if (f == "--overrelax")
            a.n_overrelax = parse_u64(get_arg(argc, argv, i, "--overrelax"), "--overrelax"), i += 2;
        else if (f == "--measure-every")
            a.measure_every = parse_u64(get_arg(argc, argv, i, "--measure-every"), "--measure-every"), i += 2;
        else if (f == "--batched") a.batched = parse_u64(get_arg(argc, argv, i, "--batched"), "--batched"), i += 2;
        else if (f == "--renorm-every")
            a.renorm_every = parse_u64(get_arg(argc, argv, i, "--renorm-every"), "--renorm-every"), i += 2;
        else if (f == "--binary") a.binary = parse_u64(get_arg(argc, argv, i, "--binary"), "--binary"), i += 2;
        else if (f == "--algorithm") a.algorithm = get_arg(argc, argv, i, "--algorithm"), i += 2;
        else i += 1;
    }

    std::vector<int> sizes;  // gpu_fss.rs:98-101: unparsable entries are dropped
    {
        size_t st = 0;
        while (st <= a.sizes_str.size()) {
            size_t e = a.sizes_str.find(',', st);
            if (e == std::string::npos) e = a.sizes_str.size();
            std::string tok = a.sizes_str.substr(st, e - st);
            const size_t b = tok.find_first_not_of(" \t\n\r"), l = tok.find_last_not_of(" \t\n\r");
            tok = b == std::string::npos ? "" : tok.substr(b, l - b + 1);
            bool ok = !tok.empty();
            for (size_t k = (tok.size() && tok[0] == '+') ? 1 : 0; k < tok.size(); ++k)
                if (tok[k] < '0' || tok[k] > '9') ok = false;
            if (tok == "+") ok = false;
            if (ok) sizes.push_back((int)std::strtoull(tok.c_str(), nullptr, 10));
            st = e + 1;
        }
    }

    validate_temp_range(a.t_min, a.t_max);  // gpu_fss.rs:103-110
    validate_samples(a.samples, "--samples");
    validate_samples(a.warmup, "--warmup");
    if (a.measure_every == 0) die("Error: --measure-every must be at least 1");
    if (a.exchange_every == 0) die("Error: --exchange-every must be at least 1");
    if (a.replicas == 0 || a.replicas > 65535) die("Error: --replicas must be in 1..65535");
    for (int L : sizes)
        if (L < 2 || L > 2046 || (L % 2) != 0) die("Error: --sizes entries must be even, 2..2046 (got " + std::to_string(L) + ")");

    mkdirs(a.outdir);
    std::string sl = "[";
    for (size_t k = 0; k < sizes.size(); ++k) sl += (k ? ", " : "") + std::to_string(sizes[k]);
    sl += "]";
    std::fprintf(stderr, "GPU FSS: model=%s, sizes=%s, T=%s..%s, replicas=%llu\n", a.model.c_str(), sl.c_str(),
                 show(a.t_min).c_str(), show(a.t_max).c_str(), a.replicas);

    int model;
    if (a.model == "ising") model = ISING_MODEL_ISING;
    else if (a.model == "xy") model = ISING_MODEL_XY;
    else if (a.model == "heisenberg") model = ISING_MODEL_HEISENBERG;
    else die("Error: --model must be ising, xy, or heisenberg");
    int cluster = -1;
    if (a.algorithm == "msc") {
        if (model != ISING_MODEL_ISING) die("Error: --algorithm msc needs --model ising");
        model = ISING_MODEL_ISING_MSC;
    } else if (a.algorithm == "sw") {
        cluster = ISING_CLUSTER_SW;
    } else if (a.algorithm == "wolff") {
        cluster = ISING_CLUSTER_WOLFF;
    } else if (a.algorithm != "metropolis") {
        die("Error: --algorithm must be metropolis, msc, sw, or wolff");
    }
    if (a.algorithm != "metropolis") std::fprintf(stderr, "  algorithm: %s\n", a.algorithm.c_str());
    if (sizes.empty()) return 0;

    if (a.batched) run(a, sizes, model, a.model.c_str(), cluster);
    else
        for (int L : sizes) run(a, std::vector<int>{L}, model, a.model.c_str(), cluster);
    return 0;
}
