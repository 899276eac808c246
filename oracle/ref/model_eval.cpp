// Driver around the reference's exported ridge models
// (ridge/out/{merbit,csr}/sampled/rd_sampled_*_model.hpp), built by
// oracle/Makefile into oracle/_ref/model_eval.  Reads lines of 10 features on
// stdin and prints "merbit csr" speedups with 17 significant digits.
#include <array>
#include <cstdio>
#include "rd_sampled_merbit_model.hpp"
#include "rd_sampled_csr_model.hpp"

int main()
{
    std::array<double, 10> f{};
    while (true) {
        for (int i = 0; i < 10; ++i)
            if (std::scanf("%lf", &f[i]) != 1) return 0;
        std::printf("%.17g %.17g\n", rd_sampled_merbit_model::PredictSpeedup(f),
                    rd_sampled_csr_model::PredictSpeedup(f));
    }
}
