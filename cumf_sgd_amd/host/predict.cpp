// cumf_predict: the reference's evaluator (test/mf-predict.cpp) on the GPU.
//   cumf_predict [-e eval] [--gpu N] test_file model_file [output_file]
// Reads the same 12-byte test triplets and model file, evaluates the
// criterion with the device kernel and prints the same result line
// ("RMSE = 0.1234", test/mf-predict.cpp:129-182).  Criteria 0 RMSE, 1 MAE,
// 2 GKL, 5 LOGLOSS, 6 ACC, 10/11 row/column MPR, 12/13 row/column AUC.
// output_file is accepted and, as in the reference, not written.
// Exit codes: 1 usage, 2 read, 3 evaluation.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/cumf_sgd_amd.h"

namespace {

const char* kHelp =
    "usage: cumf_predict [options] test_file model_file [output_file]\n"
    "\n"
    "options:\n"
    "-e <eval>: specify the evaluation criterion (default 0)\n"
    "\t 0 -- root mean square error\n"
    "\t 1 -- mean absolute error\n"
    "\t 2 -- generalized KL-divergence\n"
    "\t 5 -- logarithmic error\n"
    "\t 6 -- accuracy\n"
    "\t10 -- row-wise mean percentile rank\n"
    "\t11 -- column-wise mean percentile rank\n"
    "\t12 -- row-wise area under the curve\n"
    "\t13 -- column-wise area under the curve\n"
    "--gpu <id>: device (default 0)\n";

struct Args {
    int eval = CUMF_EVAL_RMSE;
    int gpu = 0;
    std::string test_path, model_path;
};

Args parse(int argc, char** argv)
{
    if (argc == 1) throw std::invalid_argument(kHelp);
    Args a;
    int i = 1;
    for (; i < argc; ++i) {
        if (!std::strcmp(argv[i], "-e")) {
            if (i + 1 >= argc) throw std::invalid_argument("need to specify evaluation criterion after -e");
            a.eval = std::atoi(argv[++i]);
            if (a.eval != CUMF_EVAL_RMSE && a.eval != CUMF_EVAL_MAE && a.eval != CUMF_EVAL_GKL &&
                a.eval != CUMF_EVAL_LOGLOSS && a.eval != CUMF_EVAL_ACC && (a.eval < 10 || a.eval > 13))
                throw std::invalid_argument("unknown evaluation criterion");
        } else if (!std::strcmp(argv[i], "--gpu")) {
            if (i + 1 >= argc) throw std::invalid_argument("need to specify a device after --gpu");
            a.gpu = std::atoi(argv[++i]);
        } else {
            break;
        }
    }
    if (i + 2 > argc) throw std::invalid_argument("cannot find test or model file");
    if (i + 3 < argc) throw std::invalid_argument("too many arguments");
    a.test_path = argv[i];
    a.model_path = argv[i + 1];
    return a;
}

const char* label(int eval)
{
    switch (eval) {
    case CUMF_EVAL_MAE: return "MAE";
    case CUMF_EVAL_GKL: return "GKL";
    case CUMF_EVAL_LOGLOSS: return "LOGLOSS";
    case CUMF_EVAL_ACC: return "ACCURACY";
    case CUMF_EVAL_ROW_MPR: return "Row-wise MPR";
    case CUMF_EVAL_COL_MPR: return "Column-wise MPR";
    case CUMF_EVAL_ROW_AUC: return "Row-wise AUC";
    case CUMF_EVAL_COL_AUC: return "Colmn-wise AUC";   // the reference's spelling (mf-predict.cpp:168)
    default: return "RMSE";
    }
}

}  // namespace

int main(int argc, char** argv)
{
    Args a;
    try {
        a = parse(argc, argv);
    } catch (const std::invalid_argument& e) {
        std::printf("%s\n", e.what());
        return 1;
    }
    int32_t m = 0, n = 0, k = 0;
    float b = 0.0f;
    float *P = nullptr, *Q = nullptr;
    if (cumf_load_model(a.model_path.c_str(), &m, &n, &k, &b, &P, &Q) != CUMF_OK) {
        std::fprintf(stderr, "cannot load model from %s: %s\n", a.model_path.c_str(), cumf_last_error());
        return 2;
    }
    cumf_node* R = nullptr;
    int64_t nnz = 0;
    int32_t tm = 0, tn = 0;
    if (cumf_read_problem(a.test_path.c_str(), &R, &nnz, &tm, &tn) != CUMF_OK) {
        std::fprintf(stderr, "cannot read %s: %s\n", a.test_path.c_str(), cumf_last_error());
        cumf_free_host(P);
        cumf_free_host(Q);
        return 2;
    }
    cumf_ctx* c = nullptr;
    double value = 0.0;
    int rc = cumf_create(a.gpu, &c);
    if (rc == CUMF_OK) rc = cumf_set_model(c, m, n, k, P, Q);
    if (rc == CUMF_OK) rc = cumf_set_test(c, R, nnz);
    if (rc == CUMF_OK) rc = cumf_test_eval(c, a.eval, 1.0f, b, &value);
    if (rc != CUMF_OK) std::fprintf(stderr, "evaluation failed: %s\n", cumf_last_error());
    if (c) cumf_destroy(c);
    cumf_free_host(R);
    cumf_free_host(P);
    cumf_free_host(Q);
    if (rc != CUMF_OK) return 3;
    std::printf("%s = %.4f\n", label(a.eval), value);
    return 0;
}
