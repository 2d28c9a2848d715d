/* Translation unit that instantiates the reference evaluator (src/hand_eval.h,
 * included from the reference tree where it lies) and exports it, so tests can
 * write golden vectors from it. Test infrastructure only. */
#include "hand_eval.h"

uint32_t ref_eval7(const int cards[7]) { return eval7(cards); }
uint32_t ref_eval5(int c0, int c1, int c2, int c3, int c4) { return eval5(c0, c1, c2, c3, c4); }
