N_cell: 100
N_dir: 8
N_band: 1
num_theta: 8
num_phi: 1

Iteration #0	 Margin per band per cell: 4.14412e+16
Iteration #1	 Margin per band per cell: 3.96596e+14
Iteration #2	 Margin per band per cell: 3.14363e+12
Iteration #3	 Margin per band per cell: 2.4814e+10
Iteration #4	 Margin per band per cell: 1.95847e+08
Iteration #5	 Margin per band per cell: 1.54575e+06
Iteration #6	 Margin per band per cell: 12200
Iteration #7	 Margin per band per cell: 96.2899
Iteration #8	 Margin per band per cell: 0.759979
Iteration #9	 Margin per band per cell: 0.00599821
Iteration #10	 Margin per band per cell: 4.73443e-05

heat flux of Boundary Condition #0: -2.62451e+10
heat flux of Boundary Condition #1: 2.62451e+10
