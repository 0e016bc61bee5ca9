/* Prints the C layout of the public structs so the Python binding can be checked against it. */
#include <stddef.h>
#include <stdio.h>

#include "dyc.h"

int main(void)
{
    printf("config %zu %zu %zu %zu %zu %zu %zu\n", sizeof(dyc_config), offsetof(dyc_config, lower_bound),
           offsetof(dyc_config, auto_resize), offsetof(dyc_config, max_blocks), offsetof(dyc_config, min_capacity),
           offsetof(dyc_config, insert_order), offsetof(dyc_config, region_mib));
    printf("stats %zu %zu %zu %zu %zu\n", sizeof(dyc_stats), offsetof(dyc_stats, buckets),
           offsetof(dyc_stats, pending), offsetof(dyc_stats, rejected), offsetof(dyc_stats, lost));
    return 0;
}
