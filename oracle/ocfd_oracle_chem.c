/* ocfd_oracle_chem.c -- CPU restatement of the chemistry path (test
 * infrastructure only; see ocfd_oracle.c). */
int orc_chem_placeholder(void) { return 0; }
