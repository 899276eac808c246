/*
 * predict.c -- oracle restatement of the RD headroom predictor.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   model 0 (MERBIT): include/model.cuh:91-94 == ridge/out/merbit/sampled/
 *                     merbit_sampled_deploy_train_all_aug_model.json
 *                     "ridge_raw_finite_features" (speedup = exp(b + sum w_i r_i))
 *   model 1 (CSR):    ridge/out/csr/sampled/csr_sampled_deploy_train_all_aug_model.json
 *   *_preprocessed:   model.cuh:96-105 (median imputer, standard scaler, scaled
 *                     ridge) for feature vectors with missing entries.
 * The constants are data of the fitted models, transcribed from those JSON
 * files; the tests re-read the JSON fixtures and compare digit for digit.
 *
 * Pinned by: the JSON fixtures (tests/golden/model_*.json, copied from
 * ridge/out) and the reference headers compiled into oracle/_ref evaluated on
 * random feature vectors (tests/golden/predict_ref.json).
 */
#include "oracle.h"
#include <math.h>

static const double kMerbitRawIntercept = 0.26649956401604796;
static const double kMerbitRawCoef[10] = {
    -0.3362573716950983, -0.2557862191260704, -0.24248079087236038, -0.0603012191232085,
    0.28053687310764397, -0.016399984212930544, -0.04634379957053418, -0.30165791361557115,
    2.754071221226108, 2.3042879761742348};

static const double kCsrRawIntercept = 0.256523531597976584;
static const double kCsrRawCoef[10] = {
    -3.23804724610474881e-01, -3.00486639537123124e-01, 5.97989718749250988e-03,
    -6.04430805086655781e-02, 4.52291486900478046e-01, -1.28875868013209188e-01,
    2.97368471773662213e-01, -4.54258534668076253e-01, 1.76762766759557954e+00,
    2.45978651843712059e+00};

/* MERBIT preprocessing variant (model.cuh:10-65) */
static const double kMerbitMedian[10] = {
    9.77429499999999951e-02, 1.15494500000000000e-01, 2.74992000000000014e-02,
    2.36329000000000011e-01, 8.61308000000000074e-02, 6.44094500000000070e-02,
    5.81218500000000009e-03, 0.0, 0.0, 0.0};
static const double kMerbitMean[10] = {
    1.76441841357901252e-01, 1.66786440354320997e-01, 5.14436632791358053e-02,
    2.49631313800000004e-01, 1.07480775665555556e-01, 1.07620488881728413e-01,
    7.15868577962963104e-02, 3.38007868471604939e-02, 2.02931203872839527e-02,
    1.49147190641975309e-02};
static const double kMerbitScale[10] = {
    1.94064710365436016e-01, 1.67556299251633567e-01, 7.21580093552929153e-02,
    1.77010135442985606e-01, 1.02656756737544291e-01, 1.16636224356467907e-01,
    1.09343083745194233e-01, 8.21079892355034563e-02, 6.79355186130273514e-02,
    6.63546719601208340e-02};
static const double kMerbitScaledCoef[10] = {
    -6.52556894462520132e-02, -4.28585922763317709e-02, -1.74969311762466044e-02,
    -1.06739269643762909e-02, 2.87990055385227359e-02, -1.91283223810189860e-03,
    -5.06737395751141566e-03, -2.47685247239517459e-02, 1.87099256711209222e-01,
    1.52900272760692069e-01};
static const double kMerbitScaledIntercept = 2.42110819725922966e-01;

double oracle_predict(const double* feat, int32_t model)
{
    const double* w = model == 1 ? kCsrRawCoef : kMerbitRawCoef;
    double v = model == 1 ? kCsrRawIntercept : kMerbitRawIntercept;
    for (int i = 0; i < 10; ++i) v += w[i] * feat[i];
    return exp(v);
}

double oracle_predict_preprocessed(const double* feat, const uint8_t* missing, int32_t model)
{
    (void)model;  /* only the MERBIT header ships the preprocessing variant */
    double v = kMerbitScaledIntercept;
    for (int i = 0; i < 10; ++i) {
        const double x = missing[i] ? kMerbitMedian[i] : feat[i];
        v += kMerbitScaledCoef[i] * ((x - kMerbitMean[i]) / kMerbitScale[i]);
    }
    return exp(v);
}
