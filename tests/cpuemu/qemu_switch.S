/* Stack switch of the CPU emulation's workgroup-thread fibers (x86-64 SysV;
 * TEST INFRASTRUCTURE ONLY).  qemu_switch(void **save, void *to): push the
 * callee-saved registers, store the stack pointer in *save, load `to`, pop
 * the registers saved there and return into that stack. */
    .text
    .globl qemu_switch
    .type qemu_switch, @function
qemu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size qemu_switch, .-qemu_switch
    .section .note.GNU-stack,"",@progbits
