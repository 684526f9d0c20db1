/* Links the gfx950 code objects built by hipcc --genco into the host
 * library as read-only data.  The build passes each path as a macro. */
    .section .rodata
    .p2align 12
    .globl psg_sieve_hsaco_start
    .type psg_sieve_hsaco_start, @object
psg_sieve_hsaco_start:
    .incbin PSG_SIEVE_HSACO
    .globl psg_sieve_hsaco_end
psg_sieve_hsaco_end:
    .byte 0
    .p2align 12
    .globl psg_bench_hsaco_start
    .type psg_bench_hsaco_start, @object
psg_bench_hsaco_start:
    .incbin PSG_BENCH_HSACO
    .globl psg_bench_hsaco_end
psg_bench_hsaco_end:
    .byte 0
    .p2align 12
    .globl psg_mlp_hsaco_start
    .type psg_mlp_hsaco_start, @object
psg_mlp_hsaco_start:
    .incbin PSG_MLP_HSACO
    .globl psg_mlp_hsaco_end
psg_mlp_hsaco_end:
    .byte 0
    .section .note.GNU-stack,"",@progbits
