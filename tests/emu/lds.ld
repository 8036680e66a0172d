/* Gathers the function-local statics of namespace jam (the kernels'
   __shared__ arrays under the emulation) into one section, so that the
   emulator can poison all of them before each workgroup. */
SECTIONS
{
  jam_lds : { PROVIDE(__start_jam_lds = .); *(.bss._ZZN3jam*) PROVIDE(__stop_jam_lds = .); }
}
INSERT AFTER .bss;
