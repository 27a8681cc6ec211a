// cumf_sgd command line: the reference's singleGPU CLI (singleGPU/main.cpp)
// on top of the C ABI.  Flags, defaults and quirks follow
// singleGPU/main.cpp:33-228 and Parameter() (singleGPU/sgd.h:45):
//   -g gpu -l lambda[,lambda_q] -k k -t iters -r lrate -a alpha -b beta
//   -s workers (rounded up to a multiple of 4) -u -v -x -y grid
//   -p test_file  train_file [model_file]
// Additions (long options, no reference counterpart):
//   --seed N        Philox key and srand seed (reference: time(NULL))
//   --precision P   fp32 | fp16 (default, the reference's) | bf16 | bf16sr | fp8
//   --serial        one worker, deterministic window order
//   --schedule S    hogwild | grid | stratum | tile (default: a 1x1 grid trains
//                   the library's choice for the ratings, cumf_default_schedule;
//                   a larger grid trains its blocks in the reference's order)
//   --lanes G       lanes per update (8..64)
//   --rounds R      stratum schedule (implied on a 1x1 grid): passes over the
//                   diagonals per epoch
//                   (s comes from a square -u/-v grid, else 24)
//   --coherent      device-scope coherent factor accesses
// With -p the test RMSE is printed after every epoch.  A 1x1 grid streams the
// files straight into device memory (cumf_sgd_train_file); gridded modes keep
// the reference's host shuffle and cycle-swap partition (cumf_sgd_train).
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cumf_sgd_amd.h"

namespace {

struct Args {
    cumf_param p;
    std::string tr_path, va_path, model_path;
    uint64_t seed = 1;
    int precision = CUMF_FP16;
    int schedule = -1;
    int lanes = 0;
    int coherent = 0;
    int rounds = 0;
};

bool is_numerical(const char* s)
{
    int c = 0;
    for (; *s; ++s)
        if (std::isdigit((unsigned char)*s)) ++c;
    return c > 0;
}

Args parse(int argc, char** argv)
{
    if (argc == 1) throw std::invalid_argument("error");
    Args a;
    cumf_param& p = a.p;
    std::memset(&p, 0, sizeof(p));
    p.k = 80; p.num_workers = 12; p.u_grid = p.v_grid = p.x_grid = p.y_grid = 1;
    p.num_iters = 30; p.gpu = 3; p.lambda_p = p.lambda_q = 0.05f; p.lrate = 0.0f;
    p.alpha = 0.01f; p.beta = 0.01f;
    int i;
    auto need = [&](const char* what) {
        if (i + 1 >= argc) throw std::invalid_argument(std::string("need to specify ") + what);
        ++i;
    };
    auto num = [&](const char* flag) {
        if (!is_numerical(argv[i])) throw std::invalid_argument(std::string(flag) + " should be followed by a number");
    };
    for (i = 1; i < argc; ++i) {
        std::string s = argv[i];
        if (s == "-g") { need("the id of GPUs after -g"); num("-g"); p.gpu = std::atoi(argv[i]); }
        else if (s == "-l") {
            need("lambda after -l");
            char* pch = std::strtok(argv[i], ",");
            if (!pch || !is_numerical(pch)) throw std::invalid_argument("regularization coefficient should be a number");
            p.lambda_p = p.lambda_q = (float)std::strtod(pch, nullptr);
            pch = std::strtok(nullptr, ",");
            if (pch) {
                if (!is_numerical(pch)) throw std::invalid_argument("regularization coefficient should be a number");
                p.lambda_q = (float)std::strtod(pch, nullptr);
            }
        }
        else if (s == "-k") { need("number of factors after -k"); num("-k"); p.k = std::atoi(argv[i]); }
        else if (s == "-t") { need("number of iterations after -t"); num("-i"); p.num_iters = std::atoi(argv[i]); }
        else if (s == "-r") { need("eta after -r"); num("-r"); p.lrate = (float)std::atof(argv[i]); }
        else if (s == "-a") { need("eta after -a"); num("-a"); p.alpha = (float)std::atof(argv[i]); }
        else if (s == "-b") { need("eta after -b"); num("-b"); p.beta = (float)std::atof(argv[i]); }
        else if (s == "-s") {
            need("number of parallel workers after -s(multiples of 4)");
            num("-s");
            p.num_workers = ((std::atoi(argv[i]) + 3) / 4) * 4;   // singleGPU/main.cpp:139
        }
        else if (s == "-u") { need("number of u grid after -u"); num("-u"); p.u_grid = std::atoi(argv[i]); }
        else if (s == "-v") { need("number of v grid after -v"); num("-v"); p.v_grid = std::atoi(argv[i]); }
        else if (s == "-x") { need("number of x grid after -x"); num("-x"); p.x_grid = std::atoi(argv[i]); }
        else if (s == "-y") { need("number of y grid after -y"); num("-y"); p.y_grid = std::atoi(argv[i]); }
        else if (s == "-p") { need("path after -p"); a.va_path = argv[i]; }
        else if (s == "--seed") { need("a number after --seed"); num("--seed"); a.seed = std::strtoull(argv[i], nullptr, 10); }
        else if (s == "--precision") {
            need("fp32|fp16|bf16|bf16sr|fp8 after --precision");
            std::string v = argv[i];
            if (v == "fp32") a.precision = CUMF_FP32;
            else if (v == "fp16") a.precision = CUMF_FP16;
            else if (v == "bf16") a.precision = CUMF_BF16;
            else if (v == "fp8") a.precision = CUMF_FP8_E4M3;
            else if (v == "bf16sr") a.precision = CUMF_BF16_SR;
            else throw std::invalid_argument("unknown precision " + v);
        }
        else if (s == "--serial") a.schedule = CUMF_SCHED_SERIAL;
        else if (s == "--schedule") {
            need("hogwild|grid|stratum|tile after --schedule");
            std::string v = argv[i];
            if (v == "hogwild") a.schedule = CUMF_SCHED_HOGWILD;
            else if (v == "grid") a.schedule = CUMF_SCHED_GRID;
            else if (v == "stratum") a.schedule = CUMF_SCHED_STRATUM;
            else if (v == "tile") a.schedule = CUMF_SCHED_TILE;
            else throw std::invalid_argument("unknown schedule " + v);
        }
        else if (s == "--lanes") { need("a number after --lanes"); num("--lanes"); a.lanes = std::atoi(argv[i]); }
        else if (s == "--rounds") { need("a number after --rounds"); num("--rounds"); a.rounds = std::atoi(argv[i]); }
        else if (s == "--coherent") a.coherent = 1;
        else break;
    }
    if (i >= argc) throw std::invalid_argument("training data not specified");
    a.tr_path = argv[i++];
    if (i < argc) {
        a.model_path = argv[i];
    } else {
        const char* ptr = std::strrchr(a.tr_path.c_str(), '/');
        ptr = ptr ? ptr + 1 : a.tr_path.c_str();
        a.model_path = std::string(ptr) + ".model";
    }
    if (p.u_grid * p.v_grid == 1) p.x_grid = p.y_grid = 1;   // singleGPU/main.cpp:219-223
    p.ux = p.u_grid * p.x_grid;
    p.vy = p.v_grid * p.y_grid;
    return a;
}

void print_args(const Args& a)
{
    const cumf_param& p = a.p;
    std::printf("k          :%d\n", p.k);
    std::printf("num_workers:%d\n", p.num_workers);
    std::printf("u_grid     :%d\n", p.u_grid);
    std::printf("v_grid     :%d\n", p.v_grid);
    std::printf("x_grid     :%d\n", p.x_grid);
    std::printf("y_grid     :%d\n", p.y_grid);
    std::printf("ux         :%d\n", p.ux);
    std::printf("vy         :%d\n", p.vy);
    std::printf("num_iters  :%d\n", p.num_iters);
    std::printf("gpu        :%d\n", p.gpu);
    std::printf("lambda_p   :%.5f\n", p.lambda_p);
    std::printf("lambda_q   :%.5f\n", p.lambda_q);
    std::printf("lrate      :%.5f\n", p.lrate);
    std::printf("alpha      :%.4f\n", p.alpha);
    std::printf("beta       :%.4f\n", p.beta);
    std::printf("tr_path    :%s\n", a.tr_path.c_str());
    std::printf("va_path    :%s\n", a.va_path.c_str());
    std::printf("model_path :%s\n", a.model_path.c_str());
}

}  // namespace

int main(int argc, char** argv)
{
    Args a;
    try {
        a = parse(argc, argv);
        print_args(a);
        std::fflush(stdout);
    } catch (std::invalid_argument& e) {
        std::printf("%s\n", e.what());
        return 1;
    }
    cumf_train_opts o;
    std::memset(&o, 0, sizeof(o));
    // without --schedule a 1x1 grid trains the library's default for the
    // ratings read (cumf_default_schedule); gridded runs keep the reference's
    // block order
    o.schedule = a.schedule >= 0 ? a.schedule : CUMF_SCHED_AUTO;
    if (a.schedule < 0 && a.rounds > 0 && a.p.ux * a.p.vy == 1) o.schedule = CUMF_SCHED_STRATUM;
    o.lanes_per_sample = a.lanes;
    o.coherent = a.coherent;
    o.stratum_rounds = a.rounds;
    o.use_graph = 1;   // each epoch's launches replay as one captured graph
    std::vector<double> rmse((size_t)(a.p.num_iters > 0 ? a.p.num_iters : 1), 0.0);
    float b = 0.0f;
    if (a.p.ux * a.p.vy == 1) {
        // 1x1 grid: the files stream straight into HBM, everything else on the device
        float *P = nullptr, *Q = nullptr;
        int32_t m = 0, n = 0;
        int64_t nnz = 0;
        const bool test = !a.va_path.empty();
        int rc = cumf_sgd_train_file(&a.p, &o, a.tr_path.c_str(), test ? a.va_path.c_str() : nullptr, a.precision,
                                     a.seed, &P, &Q, &m, &n, &nnz, &b, test ? rmse.data() : nullptr);
        if (rc != CUMF_OK) {
            std::fprintf(stderr, "%s%s\n", rc == CUMF_ERR_IO ? "" : "training failed: ", cumf_last_error());
            return rc == CUMF_ERR_IO ? 2 : 3;
        }
        std::printf("m:%d, n:%d, nnz:%lld\n", m, n, (long long)nnz);
        if (test)
            for (int ep = 0; ep < a.p.num_iters; ++ep) std::printf("iter %3d  test RMSE %.4f\n", ep, rmse[ep]);
        std::printf("save_model() is called\n");
        rc = cumf_save_model(a.model_path.c_str(), m, n, a.p.k, b, P, Q);
        cumf_free_host(P);
        cumf_free_host(Q);
        if (rc != CUMF_OK) {
            std::fprintf(stderr, "%s\n", cumf_last_error());
            return 4;
        }
        std::printf("\ntraining application finished...\n\n\n");
        return 0;
    }
    cumf_node* R = nullptr;
    int64_t nnz = 0;
    int32_t m = 0, n = 0;
    if (cumf_read_problem(a.tr_path.c_str(), &R, &nnz, &m, &n) != CUMF_OK) {
        std::fprintf(stderr, "%s\n", cumf_last_error());
        return 2;
    }
    std::printf("m:%d, n:%d, nnz:%lld\n", m, n, (long long)nnz);
    cumf_node* T = nullptr;
    int64_t tnnz = 0;
    int32_t tm = 0, tn = 0;
    if (!a.va_path.empty() && cumf_read_problem(a.va_path.c_str(), &T, &tnnz, &tm, &tn) != CUMF_OK) {
        std::fprintf(stderr, "%s\n", cumf_last_error());
        cumf_free_host(R);
        return 2;
    }
    std::vector<float> P((size_t)m * a.p.k), Q((size_t)n * a.p.k);
    int rc = cumf_sgd_train(&a.p, &o, R, nnz, m, n, a.precision, a.seed, T, tnnz, P.data(), Q.data(), &b,
                            T ? rmse.data() : nullptr);
    if (rc != CUMF_OK) {
        std::fprintf(stderr, "training failed: %s\n", cumf_last_error());
        cumf_free_host(R);
        cumf_free_host(T);
        return 3;
    }
    if (T)
        for (int ep = 0; ep < a.p.num_iters; ++ep) std::printf("iter %3d  test RMSE %.4f\n", ep, rmse[ep]);
    std::printf("save_model() is called\n");
    rc = cumf_save_model(a.model_path.c_str(), m, n, a.p.k, b, P.data(), Q.data());
    cumf_free_host(R);
    cumf_free_host(T);
    if (rc != CUMF_OK) {
        std::fprintf(stderr, "%s\n", cumf_last_error());
        return 4;
    }
    std::printf("\ntraining application finished...\n\n\n");
    return 0;
}
