"""Host-side mirror of the reference's join operator modules
(operators/partition_sort.py, operators/join_core.py, operators/hashjoin.py,
operators/partitions.py): the same names, argument meanings and error
conventions, over the HIP kernels behind include/pageflow_join.h."""
