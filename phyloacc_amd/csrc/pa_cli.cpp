// pa_cli.cpp -- drop-in command line for PhyloAcc-ST runs on the MI355X engine.
//
// Usage: phyloacc-st-gpu <param file> [--dry-run]
// The program is pa_cli_main, exported by the engine library, so that a
// host process can also run it in-process.
//
// Reads the reference's parameter file (src/PhyloAcc-ST/main.cpp:100-287),
// the .mod tree (newick.cpp:431-517, through pa_parse_mod), the FASTA
// alignment and the BED-like segment file (profile.cpp:17-126), runs the
// three model chains of every selected element through the C ABI and
// writes the reference's output files:
//   <RESULT_FOLDER>/<PREFIX>_species_names.txt       (main.cpp:378-391)
//   <PREFIX>_hyper.txt                               (main.cpp:365-369)
//   <PREFIX>_rate_postZ_M{0,1,2}.txt                 (header main.cpp:383-391, rows bpp_c2.cpp:739-785)
//   <PREFIX>_elem_lik.txt, <PREFIX>_M{0,1,2}_elem_Z.txt   (bpp.cpp:566-618)
// File M1 holds the accelerated-only model (resZ 2) and M2 the full model
// (resZ 1), as in the reference.  Keys the reference does not know:
//   PRECISION fp32|fp64   partial precision (default fp32)
//   DEVICE n              GPU ordinal (default 0)
//   BATCH_ELEMENTS n      elements per device launch (default 4096)
// --dry-run stops after loading and encoding (no device needed) and prints
// one summary line per element.
#include <dirent.h>
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/phyloacc_gpu.h"

namespace {

struct Params {
    std::string phytree, align, seg, id_file, out_dir, prefix = "test";
    std::string outgroup, target, conserve;
    int batch = -1, chain = 0, num_thread = 1;
    bool sample_hyper = false, verbose = false;
    int precision = PA_FP32, device = 0, batch_elements = 4096;
    pa_options o;
};

std::string trim(const std::string &s, const char *chars = " \t\n\r")
{
    const size_t a = s.find_first_not_of(chars);
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(chars);
    return s.substr(a, b - a + 1);
}

std::vector<std::string> split(const std::string &s, char d)
{
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, d)) out.push_back(item);
    return out;
}

bool to_bool(const std::string &s) { return s == "true" || s == "1"; }

// Input errors end the run with status 1 (the reference calls exit(1)); the
// status travels as an exception to pa_cli_main so that an in-process caller
// survives it.
struct CliExit {
    int code;
};

[[noreturn]] void die(const std::string &msg)
{
    std::cerr << msg << std::endl;
    throw CliExit{1};
}

// LoadParams (main.cpp:100-287): "KEY value" lines, '#' comments.
Params load_params(const char *path)
{
    Params P;
    pa_default_options(&P.o);
    std::ifstream in(path);
    if (!in) die(std::string("Cannot open the parameters file: ") + path);
    std::string line;
    int line_num = 1;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::vector<std::string> w;
        std::string word;
        while (ls >> word) w.push_back(word);
        if (w.empty() || w[0][0] == '#') continue;
        if (w.size() != 2)
            die("Line " + std::to_string(line_num) + " in the parameter file is not formatted correctly: " + line);
        const std::string &k = w[0], &v = w[1];
        pa_options &o = P.o;
        if (k == "PHYTREE_FILE") P.phytree = v;
        else if (k == "ALIGN_FILE") P.align = v;
        else if (k == "SEG_FILE") P.seg = v;
        else if (k == "ID_FILE") P.id_file = v;
        else if (k == "BATCH") P.batch = std::stoi(v);
        else if (k == "RESULT_FOLDER") P.out_dir = v;
        else if (k == "PREFIX") P.prefix = v;
        else if (k == "SEED") o.seed = (uint64_t)std::stoi(v);
        else if (k == "INIT_CONSERVE_RATE") o.init_crate = std::stod(v);
        else if (k == "INIT_ACCE_RATE") o.init_nrate = std::stod(v);
        else if (k == "CONSERVE_PRIOR_A") o.cprior_a = std::stod(v);
        else if (k == "CONSERVE_PRIOR_B") o.cprior_b = std::stod(v);
        else if (k == "ACCE_PRIOR_A") o.nprior_a = std::stod(v);
        else if (k == "ACCE_PRIOR_B") o.nprior_b = std::stod(v);
        else if (k == "ROPT") o.ropt = std::stoi(v);
        else if (k == "CUB") o.cub = std::stod(v);
        else if (k == "NLB") o.nlb = std::stod(v);
        else if (k == "BURNIN") o.burnin = std::stoi(v);
        else if (k == "MCMC") o.mcmc = std::stoi(v);
        else if (k == "ADAPT_FREQ") o.adapt_freq = std::stoi(v);
        else if (k == "INIT_LRATE") o.init_lrate = std::stod(v);
        else if (k == "INIT_LRATE2") o.init_lrate2 = std::stod(v);
        else if (k == "INIT_GRATE") o.init_grate = std::stod(v);
        else if (k == "HYPER_LRATE_A") o.hyper_lrate_a = std::stod(v);
        else if (k == "HYPER_LRATE_B") o.hyper_lrate_b = std::stod(v);
        else if (k == "HYPER_GRATE_A") o.hyper_grate_a = std::stod(v);
        else if (k == "HYPER_GRATE_B") o.hyper_grate_b = std::stod(v);
        else if (k == "HYPER_LRATE2_A") o.hyper_lrate2_a = std::stod(v);
        else if (k == "HYPER_LRATE2_B") o.hyper_lrate2_b = std::stod(v);
        else if (k == "CHAIN") P.chain = std::stoi(v);
        else if (k == "OUTGROUP") P.outgroup = v;
        else if (k == "TARGETSPECIES") P.target = v;
        else if (k == "CONSERVE") P.conserve = v;
        else if (k == "CONSERVE_PROP") o.conserve_prop = std::stod(v);
        else if (k == "GAP_PROP") o.gap_prop = std::stod(v);
        else if (k == "CONSTOMIS") o.constomis = std::stod(v);
        else if (k == "GAPCHAR") o.gapchar = (unsigned char)v[0];
        else if (k == "PRUNE_TREE") o.prune_tree = to_bool(v);
        else if (k == "TRIM_GAP_PERCENT") o.trim_gap_percent = std::stod(v);
        else if (k == "MIN_LEN") o.min_len = std::stoi(v);
        else if (k == "INDEL" || k == "INDEL2" || k == "SAMPLE_INDEL") { /* read and unused, as in the reference */ }
        else if (k == "SAMPLE_HYPER") P.sample_hyper = to_bool(v);
        else if (k == "VERBOSE") P.verbose = to_bool(v);
        else if (k == "NUM_THREAD") P.num_thread = std::stoi(v);
        else if (k == "PRECISION") P.precision = (v == "fp64" || v == "FP64") ? PA_FP64 : PA_FP32;
        else if (k == "DEVICE") P.device = std::stoi(v);
        else if (k == "BATCH_ELEMENTS") P.batch_elements = std::max(1, std::stoi(v));
        else std::cout << "Unknown parameter: " << k << ", skipping..." << std::endl;
        line_num++;
    }
    const char *q = " \"\t\n";
    P.phytree = trim(P.phytree, q);
    P.align = trim(P.align, q);
    P.out_dir = trim(P.out_dir, q);
    P.seg = trim(P.seg, q);
    if (P.o.hyper_lrate2_a == 0) P.o.init_lrate2 = 0;
    return P;
}

struct Profile {
    std::vector<std::string> names, X;      // species rows, lower-cased
    std::vector<std::string> elem_names;
    std::vector<int64_t> elem_len;          // end - start of each segment line
};

// LoadPhyloProfiles (profile.cpp:17-126): FASTA rows, then "name\tstart\tend" lines.
Profile load_profile(const std::string &fasta, const std::string &seg)
{
    Profile pf;
    std::ifstream in(fasta);
    if (!in) die("(Error. Cannot open the phylogenetic profile input file: " + fasta + ")");
    std::string line, cur;
    bool have = false;
    size_t G = 0;
    auto flush = [&]() {
        if (!have) return;
        for (auto &ch : cur) ch = (char)tolower((unsigned char)ch);
        if (pf.X.empty()) G = cur.size();
        else if (cur.size() != G) die("alignment rows differ in length (" + pf.names[pf.X.size()] + ")");
        pf.X.push_back(cur);
        cur.clear();
    };
    while (std::getline(in, line)) {
        line = trim(line);
        if (!line.empty() && line[0] == '>') {
            flush();
            pf.names.push_back(trim(line.substr(1)));
            have = true;
        } else {
            cur += line;
        }
    }
    flush();
    std::ifstream sg(seg);
    if (!sg) die("(Error. Cannot open the segment input file: " + seg + ")");
    while (std::getline(sg, line)) {
        line = trim(line);
        std::vector<std::string> f = split(line, '\t');
        if (f.size() < 3) break;
        pf.elem_names.push_back(f[0]);
        pf.elem_len.push_back((int64_t)atoi(f[2].c_str()) - (int64_t)atoi(f[1].c_str()));
    }
    return pf;
}

bool dir_exists(const std::string &p)
{
    if (p.empty()) return false;
    DIR *d = opendir(p.c_str());
    if (!d) return false;
    closedir(d);
    return true;
}

std::string fmt_g(double v, int prec)
{
    char b[64];
    snprintf(b, sizeof b, "%.*g", prec, v);
    return b;
}

}  // namespace

static int cli(int argc, char **argv);

extern "C" int pa_cli_main(int argc, char **argv)
{
    try {
        const int rc = cli(argc, argv);
        std::cout.flush();
        return rc;
    } catch (const CliExit &x) {
        std::cout.flush();
        return x.code;
    } catch (const std::exception &x) {
        std::cerr << "error: " << x.what() << std::endl;
        return 1;
    }
}

static int cli(int argc, char **argv)
{
    if (argc < 2) {
        std::cout << "Please specify the path of the parameter file." << std::endl;
        return 1;
    }
    const bool dry = argc > 2 && std::string(argv[2]) == "--dry-run";
    Params P = load_params(argv[1]);
    if (!dir_exists(P.out_dir)) {
        std::cout << "output path doesn't exist or empty!" << std::endl;
        return 1;
    }
    if (P.sample_hyper) die("SAMPLE_HYPER is not supported by this engine");
    Profile pf = load_profile(P.align, P.seg);

    // tree
    std::ifstream tf(P.phytree);
    if (!tf) die("Cannot open the tree file: " + P.phytree);
    std::stringstream tss;
    tss << tf.rdbuf();
    const int MAXN = 512;
    int32_t S = 0;
    std::vector<int32_t> ch(2 * MAXN), par(MAXN);
    std::vector<double> dist(MAXN), pi(4), Q(16);
    std::vector<char> nm(64 * MAXN);
    if (pa_parse_mod(tss.str().c_str(), MAXN, &S, ch.data(), par.data(), dist.data(), pi.data(), Q.data(), nm.data()) != PA_OK)
        die("cannot parse the tree file " + P.phytree);
    const int N = 2 * S - 1;
    std::vector<std::string> node_names(N);
    for (int i = 0; i < N; i++) node_names[i] = std::string(&nm[64 * i]);

    // MatchProfAndTree (bpp.cpp:133-184): reorder rows by the tree's leaves,
    // or keep the profile order when the names do not all match
    std::vector<const char *> rows(S);
    std::vector<std::string> species(S);
    bool matched = (int)pf.names.size() >= S;
    std::vector<int> reorder(S, -1);
    for (int s = 0; s < S && matched; s++) {
        auto it = std::find(pf.names.begin(), pf.names.end(), node_names[s]);
        if (it == pf.names.end()) matched = false;
        else reorder[s] = (int)(it - pf.names.begin());
    }
    if ((int)pf.X.size() < S) die("the alignment has fewer rows than the tree has species");
    for (int s = 0; s < S; s++) {
        const int r = matched ? reorder[s] : s;
        rows[s] = pf.X[r].c_str();
        species[s] = matched ? node_names[s] : pf.names[s];
    }
    if (!matched) std::cout << "The species in phylogenetic profile and tree cannot be matched literally; using the default mapping." << std::endl;

    auto ids_of = [&](const std::string &list) {
        std::vector<int32_t> out;
        for (const std::string &nme : split(trim(list), ';')) {
            auto it = std::find(species.begin(), species.end(), nme);
            if (it == species.end()) { std::cerr << "warning: species " << nme << " is not in the tree; ignored" << std::endl; continue; }
            out.push_back((int32_t)(it - species.begin()));
        }
        return out;
    };
    std::vector<int32_t> outg = ids_of(P.outgroup), cons = ids_of(P.conserve);

    // element slices: contiguous, in segment-file order (bpp.hpp:257-274)
    const int C = (int)pf.elem_names.size();
    std::vector<int64_t> start(C);
    int64_t acc = 0;
    for (int c = 0; c < C; c++) { start[c] = acc; acc += pf.elem_len[c]; }
    if (acc > (int64_t)pf.X[0].size()) die("segments run past the end of the alignment");

    // ids: ID_FILE, BATCH thirds, or all (main.cpp:406-440)
    std::vector<int> ids;
    if (P.id_file.empty()) {
        if (P.batch == -1) for (int c = 0; c < C; c++) ids.push_back(c);
        else {
            const int t = C / 3;
            for (int c = P.batch * t; c < (P.batch + 1) * t && c < C; c++) ids.push_back(c);
        }
    } else {
        std::ifstream fi(P.id_file);
        if (!fi) die("Cannot open the id file: " + P.id_file);
        std::string l;
        while (std::getline(fi, l)) {
            std::istringstream ls(l);
            std::string tok;
            ls >> tok;
            tok = trim(tok);
            if (tok.empty()) continue;
            const int c = atoi(tok.c_str());
            if (c < 0 || c >= C) die("element id " + tok + " is out of range");
            ids.push_back(c);
        }
    }
    std::cout << ids.size() << " of elements to be computed" << std::endl;

    pa_model m;
    m.S = S; m.children = ch.data(); m.parent = par.data(); m.dist = dist.data(); m.pi = pi.data(); m.Q = Q.data();
    m.outgroup = outg.data(); m.n_outgroup = (int32_t)outg.size();
    m.conserve = cons.data(); m.n_conserve = (int32_t)cons.size();
    pa_engine *eng = nullptr;
    if (pa_engine_create(&m, &P.o, P.device, &eng) != PA_OK) die(std::string("engine: ") + pa_last_error(eng));
    struct Guard {
        pa_engine *e;
        ~Guard() { pa_engine_destroy(e); }
    } guard{eng};
    std::vector<int64_t> st_ids(ids.size());
    std::vector<int32_t> ln_ids(ids.size()), eid(ids.size()), est(ids.size());
    for (size_t i = 0; i < ids.size(); i++) { st_ids[i] = start[ids[i]]; ln_ids[i] = (int32_t)pf.elem_len[ids[i]]; }
    if (pa_elements_add(eng, rows.data(), st_ids.data(), ln_ids.data(), (int32_t)ids.size(), eid.data(), est.data()) != PA_OK)
        die(std::string("encoding: ") + pa_last_error(eng));

    const std::string base = P.out_dir + "/" + P.prefix;
    if (dry) {
        for (size_t i = 0; i < ids.size(); i++) {
            pa_element_info inf;
            pa_element_get(eng, eid[i], &inf, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
            std::cout << ids[i] << "\t" << pf.elem_names[ids[i]] << "\t" << (est[i] == PA_ELEM_FILTERED ? "filtered" : "ok")
                      << "\tGG=" << inf.GG << "\tP=" << inf.P << "\troot=" << inf.root << "\tnodes=" << inf.n_nodes << "\n";
        }
        return 0;
    }

    // outputs written up front (main.cpp:360-391)
    {
        std::ofstream sp(base + "_species_names.txt");
        for (int s = 0; s < N; s++) sp << node_names[s] << "\n";
        std::ofstream hy(base + "_hyper.txt");
        hy << "iter\tnprior_a\tnprior_b\tcprior_a\tcprior_b\tprior_l_a\tprior_l_b\tprior_g_a\tprior_g_b\n";
        hy << 0 << "\t" << fmt_g(P.o.nprior_a, 6) << "\t" << fmt_g(P.o.nprior_b, 6) << "\t" << fmt_g(P.o.cprior_a, 6) << "\t"
           << fmt_g(P.o.cprior_b, 6) << "\t" << fmt_g(P.o.hyper_lrate_a, 6) << "\t" << fmt_g(P.o.hyper_lrate_b, 6) << "\t"
           << fmt_g(P.o.hyper_grate_a, 6) << "\t" << fmt_g(P.o.hyper_grate_b, 6) << "\n";
    }
    std::ofstream outZ[3];   // by model slot: M0 (resZ 0), M1 (resZ 2), M2 (resZ 1)
    for (int k = 0; k < 3; k++) {
        outZ[k].open(base + "_rate_postZ_M" + std::to_string(k) + ".txt");
        outZ[k] << "No.\tn_rate\tc_rate\tg_rate\tl_rate\tl2_rate";
        for (int s = 0; s < N; s++)
            for (int z = 0; z < 4; z++) outZ[k] << "\t" << node_names[s] << "_" << z;
        outZ[k] << "\n";
    }

    const int T = P.o.burnin + P.o.mcmc;
    const size_t nid = ids.size();
    std::vector<double> vb((size_t)C * 3, 0.0), maxll((size_t)C * 3, 0.0);
    std::vector<int8_t> maxZ((size_t)C * 3 * N, 0);
    pa_run_config cfg = {P.precision, PA_RUN_BATCHED, 0, 1};
    for (int iter = 0; iter < P.chain; iter++) {
        std::cout << "Running MCMC chain " << iter + 1 << " ..." << std::endl;
        for (size_t b0 = 0; b0 < nid; b0 += (size_t)P.batch_elements) {
            std::vector<int32_t> run;
            std::vector<size_t> pos;
            for (size_t i = b0; i < std::min(nid, b0 + (size_t)P.batch_elements); i++)
                if (est[i] == PA_ELEM_OK) { run.push_back(eid[i]); pos.push_back(i); }
            if (run.empty()) continue;
            const size_t n = run.size();
            std::vector<pa_chain_result> res(n * 3);
            std::vector<double> full(n * 3 * T), nr(n * 3 * T), cr(n * 3 * T), gr(n * 3 * T), lr(n * 3 * T), l2(n * 3 * T);
            std::vector<int8_t> Z(n * 3 * (size_t)T * N), mz(n * 3 * N);
            pa_traces tr;
            memset(&tr, 0, sizeof tr);
            tr.full = full.data(); tr.nrate = nr.data(); tr.crate = cr.data(); tr.grate = gr.data();
            tr.lrate = lr.data(); tr.l2rate = l2.data(); tr.Z = Z.data(); tr.max_Z = mz.data();
            if (pa_run_chains(eng, run.data(), (int32_t)n, &cfg, res.data(), &tr, nullptr) != PA_OK)
                die(std::string("device: ") + pa_last_error(eng));
            for (size_t i = 0; i < n; i++) {
                const int c = ids[pos[i]];
                for (int k = 0; k < 3; k++) {
                    const size_t o = (i * 3 + k) * (size_t)T;
                    double v = 0, med[5];
                    std::vector<double> postZ((size_t)N * 4);
                    if (pa_chain_summary(eng, run[i], k, &full[o], &nr[o], &cr[o], &gr[o], &lr[o], &l2[o],
                                         &Z[o * N], &mz[(i * 3 + k) * N], P.precision, &v, med, postZ.data()) != PA_OK)
                        die(std::string("summary: ") + pa_last_error(eng));
                    vb[(size_t)c * 3 + k] = v;
                    maxll[(size_t)c * 3 + k] = res[i * 3 + k].max_loglik;
                    memcpy(&maxZ[((size_t)c * 3 + k) * N], &mz[(i * 3 + k) * N], N);
                    outZ[k] << c;
                    for (int q = 0; q < 5; q++) outZ[k] << "\t" << fmt_g(med[q], 6);
                    for (int s = 0; s < N; s++)
                        for (int z = 0; z < 4; z++) outZ[k] << "\t" << fmt_g(postZ[s * 4 + z], 6);
                    outZ[k] << "\n";
                }
            }
            for (int k = 0; k < 3; k++) outZ[k].flush();
        }
        // BPP::Output_init (bpp.cpp:566-618)
        std::ofstream lik(base + "_elem_lik.txt");
        lik << "No.\tID\tloglik_Null\tloglik_Acc\tloglik_Full\tlogBF1\tlogBF2\tloglik_Max_M0\tloglik_Max_M1\tloglik_Max_M2\n";
        for (int c : ids) {
            const double nul = vb[(size_t)c * 3 + 0], acc_ = vb[(size_t)c * 3 + 1], ful = vb[(size_t)c * 3 + 2];
            lik << c << "\t" << pf.elem_names[c] << "\t" << fmt_g(nul, 8) << "\t" << fmt_g(acc_, 8) << "\t" << fmt_g(ful, 8)
                << "\t" << fmt_g(acc_ - nul, 8) << "\t" << fmt_g(acc_ - ful, 8) << "\t" << fmt_g(maxll[(size_t)c * 3 + 0], 8)
                << "\t" << fmt_g(maxll[(size_t)c * 3 + 1], 8) << "\t" << fmt_g(maxll[(size_t)c * 3 + 2], 8) << "\n";
        }
        for (int k = 0; k < 3; k++) {
            std::ofstream ez(base + "_M" + std::to_string(k) + "_elem_Z.txt");
            ez << "No.";
            for (int s = 0; s < N; s++) ez << "\t" << node_names[s];
            ez << "\n";
            for (int c : ids) {
                ez << c;
                for (int s = 0; s < N; s++) ez << "\t" << (int)maxZ[((size_t)c * 3 + k) * N + s];
                ez << "\n";
            }
        }
    }
    return 0;
}
