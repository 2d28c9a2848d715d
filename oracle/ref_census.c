/* Census of the reference evaluator over all C(52,7) hands: category counts
 * and the sum of all ranks mod 2^64. Compiled against src/hand_eval.h in
 * place (see Makefile). Test infrastructure only: writes a golden vector. */
#include <stdio.h>
#include <stdint.h>
#include "hand_eval.h"

int main(void)
{
    uint64_t cnt[10] = {0}, sum = 0;
    int c[7];
    for (c[0] = 0; c[0] < 46; c[0]++)
    for (c[1] = c[0] + 1; c[1] < 47; c[1]++)
    for (c[2] = c[1] + 1; c[2] < 48; c[2]++)
    for (c[3] = c[2] + 1; c[3] < 49; c[3]++)
    for (c[4] = c[3] + 1; c[4] < 50; c[4]++)
    for (c[5] = c[4] + 1; c[5] < 51; c[5]++)
    for (c[6] = c[5] + 1; c[6] < 52; c[6]++) {
        uint32_t v = eval7(c);
        cnt[v >> 20]++;
        sum += v;
    }
    for (int k = 0; k < 10; k++) printf("%llu ", (unsigned long long)cnt[k]);
    printf("%llu\n", (unsigned long long)sum);
    return 0;
}
