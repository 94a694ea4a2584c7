// Drives the library's step-length search (swfwi_amd/csrc/linesearch.h, the
// host-side state machine the GPU drivers use) on a scripted objective table,
// so the line-search tests can check it against hand-derived trial sequences
// without a GPU.  Built with g++ by tests/test_oracle_sequences.py.
//
// stdin:  obj1 max_alpha3 nita preserved_alpha preserved_init
//         n
//         alpha_1 obj_1  ...  alpha_n obj_n     (floats printed with %a)
// stdout: trials <alpha ...>
//         result steplen obj alpha2 alpha3 obj2 obj3 parabolic preserved_alpha
// An alpha missing from the table exits with status 3.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../swfwi_amd/csrc/linesearch.h"

int main() {
  float obj1, max3, pa;
  int nita, pinit, n;
  if (std::scanf("%a %a %d %a %d %d", &obj1, &max3, &nita, &pa, &pinit, &n) != 6) return 2;
  std::vector<float> ta(n), to(n);
  for (int i = 0; i < n; i++)
    if (std::scanf("%a %a", &ta[i], &to[i]) != 2) return 2;
  auto objective = [&](float a) {
    for (int i = 0; i < n; i++)
      if (ta[i] == a) return to[i];
    std::printf("missing %a\n", a);
    std::exit(3);
  };
  swfwi::PreservedAlpha pres;
  pres.alpha = pa;
  pres.init = pinit != 0;
  float a2, a3;
  pres.init_alpha23(max3, a2, a3);
  swfwi::RefineAlpha ra(obj1, max3, a2, a3, nita);
  std::vector<float> trials;
  while (!ra.done()) {
    std::vector<float> o;
    for (float a : ra.pending()) {
      trials.push_back(a);
      o.push_back(objective(a));
    }
    ra.feed(o.data());
  }
  float a4, o4;
  swfwi::final_step(ra, obj1, max3, a4, o4);
  pres.alpha = a4;
  std::printf("trials");
  for (float a : trials) std::printf(" %a", a);
  std::printf("\nresult %a %a %a %a %a %a %d %a\n", a4, o4, ra.ra2, ra.ra3, ra.ro2, ra.ro3, ra.parabolic, pres.alpha);
  return 0;
}
