/* Prints the C layout of the public structs so the Python binding can be checked against it. */
#include <stddef.h>
#include <stdio.h>

#include "dyc.h"

int main(void)
{
    printf("config %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(dyc_config), offsetof(dyc_config, lower_bound),
           offsetof(dyc_config, auto_resize), offsetof(dyc_config, max_blocks), offsetof(dyc_config, min_capacity),
           offsetof(dyc_config, insert_order), offsetof(dyc_config, region_mib), offsetof(dyc_config, subtables));
    printf("stats %zu %zu %zu %zu %zu %zu\n", sizeof(dyc_stats), offsetof(dyc_stats, buckets),
           offsetof(dyc_stats, pending), offsetof(dyc_stats, rejected), offsetof(dyc_stats, lost),
           offsetof(dyc_stats, subtables));
    printf("request %zu %zu %zu %zu %zu\n", sizeof(dyc_request), offsetof(dyc_request, keys),
           offsetof(dyc_request, vals), offsetof(dyc_request, out_flags), offsetof(dyc_request, n));
    printf("server_config %zu %zu %zu %zu %zu %zu %zu\n", sizeof(dyc_server_config),
           offsetof(dyc_server_config, tick_max_requests), offsetof(dyc_server_config, item_keys),
           offsetof(dyc_server_config, pool_cap), offsetof(dyc_server_config, poll_backoff),
           offsetof(dyc_server_config, resident), offsetof(dyc_server_config, resident_idle_ms));
    printf("server_stats %zu %zu %zu %zu\n", sizeof(dyc_server_stats), offsetof(dyc_server_stats, deferred_memory),
           offsetof(dyc_server_stats, in_flight), offsetof(dyc_server_stats, loop_pauses));
    printf("stats_tail %zu %zu\n", offsetof(dyc_stats, last_max_chain), offsetof(dyc_stats, last_filled));
    return 0;
}
